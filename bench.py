"""Benchmark of the LIMA_MD step on one MI355X.

Default workload (the production configuration): 10 chains x 10 compounds x 128
particles (12,800 chain particles) plus lattice solvent at 33 nm^-3 (~24k) in a
10 nm periodic box, 16 LJ types, 2 fs, neighbour rebuild every 10 steps,
Berendsen thermostat, FP32 pairs.  Synthetic seeded data, generated here.

Timing: W warmup steps (the first captures the launch graphs), then untimed steps up to the next
rebuild-cycle boundary (warmup_align_steps) and whole cycles for 50 ms while the GPU clocks rise
(warmup_ramp_steps), then exactly K timed steps.

Prints ONE JSON line: ns/day (whole job), ms/step, the roofline of the dominant
kernel (the all-pairs solvent<-chain LJ kernel, VALU FP32 bound), and the CPU
baseline (the oracle with OpenMP on this host's cores, on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_PAIR = 23          # algorithmic FP32 flops per directed LJ pair (force only)
PEAK_FP32_TFLOPS = 157.3     # MI355X published FP32 vector/matrix peak (packed); unpacked FMA rate is half
PEAK_HBM_GBS = 8000.0
RAMP_S = 0.05                # untimed stepping before the timed region, after warmup (clock ramp)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from quantom_amd import systems
    if name == "production":
        return [systems.production()], ("production: 10 chains x 10 compounds x 128 + lattice solvent 33/nm^3, 10 nm, "
                                        "16 types (compounds built within 1.6 nm, central numbering, sphere margin 1.25 nm)")
    if name == "small":
        return [systems.small()], "small: 1 chain x 2 compounds x 128 + 3000 solvent, 5 nm, 8 types"
    if name == "large":
        return [systems.large()], ("large: 10 chains x 40 compounds x 128 + lattice solvent 33/nm^3, 18 nm, 16 types "
                                   "(compounds built within 1.6 nm, central numbering, sphere margin 1.25 nm)")
    if name == "t4lysozyme":
        from quantom_amd.abi import LimaSystem
        s = LimaSystem.load(os.path.join(ROOT, "tests", "golden", "t4lysozyme.npz"))
        return [s], ("t4lysozyme: the reference's demo (Demo/6lzm) as ingested: 36 compounds, 1293 heavy atoms, "
                     "5000 waters, 7 nm, 19 types, dt 0.1 fs")
    if name == "replicas":
        return systems.replicas(256), ("replicas: 256 systems, chain 128-1024 in compounds of 16 (box < 4.6 nm) "
                                       "or 32, solvent 500-6000, box edge at 33/nm^3 within 3-6 nm, T0 280-320 K")
    raise SystemExit(f"unknown workload {name}")


# The reference's published step time (README.md:19), quoted without naming the system;
# the shipped demo (Environment.cpp:24-33 loads Demo/6lzm) is the system it implies.
PUBLISHED_MS_PER_STEP = 3.34


def run_params(name, accumulation):
    from quantom_amd.abi import ACCUM_FIXED64, ACCUM_FP32, THERMOSTAT_REFERENCE, RunParams
    acc = ACCUM_FIXED64 if accumulation == "fixed64" else ACCUM_FP32
    if name == "t4lysozyme":
        # the reference's constants: CUTOFF 0.7 nm (Constants.cuh:52), lists every 200 steps
        # (Constants.cuh:71), the thermostat cadence of Engine::hostMaster (Engine.cu:49-75)
        return RunParams(cutoff=0.7, nlist_interval=200, thermostat=THERMOSTAT_REFERENCE, accumulation=acc,
                         max_solvent_neighbors=1024, max_compound_solvents=4096)
    params = RunParams(nlist_interval=20 if name == "large" else 10, accumulation=acc)
    if name == "large":
        params.max_compound_solvents = 32768     # compounds confined to 0.38 L span up to ~17k list entries
    return params


def demo_comparison(steps, warmup):
    """ms/step on the reference's demo system, next to the reference's published figure."""
    from quantom_amd.engine import Engine
    syss, desc = workload("t4lysozyme")
    eng = Engine(syss, run_params("t4lysozyme", "fp32"))
    eng.autotune(10)
    eng.step(warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    eng.close()
    return {"workload": desc, "steps": steps, "ms_per_step": ms, "ns_per_day": syss[0].dt * 1e6 * 86.4 / ms,
            "published_ms_per_step": PUBLISHED_MS_PER_STEP, "published_over_this": PUBLISHED_MS_PER_STEP / ms,
            "note": "published figure: unnamed NVIDIA GPU (sm_61/sm_75 build), system inferred from the shipped demo"}


def read_traffic(kernel_tag):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(systems_, params, budget_s):
    """The oracle (FP32, OpenMP on all host threads) on a bounded sample of the same system: three
    windows of steps without a rebuild (median of the window means) plus one rebuild timed alone and
    amortised over the rebuild interval."""
    import statistics

    from oracle.oracle import Oracle, num_threads
    sysm = systems_[0]
    iv = max(1, params.nlist_interval)
    o = Oracle(sysm, params, precision=32, fast=True)
    t0 = time.perf_counter()
    o.rebuild_neighbors()
    t_reb = time.perf_counter() - t0
    o.step(1)                                        # warm
    windows, steps = [], 0
    t_start = time.perf_counter()
    while len(windows) < 3:
        # steps that do not end on a rebuild (the oracle rebuilds after step k when k % iv == 0)
        n = 0
        t = 0.0
        while n < max(1, iv - 1):
            if (o.step_count + 1) % iv == 0:
                o.step(1)                            # rebuild step: not part of the windows
                continue
            t1 = time.perf_counter()
            o.step(1)
            t += time.perf_counter() - t1
            n += 1
            if time.perf_counter() - t_start > budget_s * (len(windows) + 1) / 3.0:
                break
        windows.append(t / n)
        steps += n
    ms = 1000.0 * (statistics.median(windows) + t_reb / iv)
    return {"value": 172.8 * (sysm.dt / 2e-6) / ms, "unit": "ns/day", "ms_per_step": ms,
            "cores": num_threads(True), "nproc": os.cpu_count(), "cpu_model": cpu_model(), "kind": "port",
            "sample": f"3 windows ({steps} steps without a rebuild, median of window means) plus one "
                      f"neighbour rebuild ({1000 * t_reb:.1f} ms) amortised over {iv} steps; "
                      f"oracle/liblima_oracle_fast.so (g++ -O3 -fopenmp, FMA allowed), same system"}


def sequential_baseline(syss, params, steps):
    """The same replicas stepped one after another, each in its own single-system engine (plain
    launches of one system): one engine exists at a time; its warmup and creation are untimed,
    and the timed steps of all replicas are summed."""
    from quantom_amd.engine import Engine
    total = 0.0
    for r, s in enumerate(syss):
        e = Engine([s], params)
        e.step(params.nlist_interval)
        e.sync()
        t0 = time.perf_counter()
        e.step(steps)
        total += time.perf_counter() - t0
        e.close()
        if (r + 1) % 64 == 0:
            log(f"sequential baseline: {r + 1} of {len(syss)} replicas")
    return 1000.0 * total / steps


def accuracy_report(args):
    """rms force error vs FP64 (FP32 and int64 accumulation) on the production system, and
    energy drift of a short NVE run on the small system beside the FP64 oracle's drift."""
    import numpy as np

    from oracle.oracle import KIND_CC, KIND_CS, KIND_SC, KIND_SS, Oracle
    from quantom_amd import systems
    from quantom_amd.abi import ACCUM_FIXED64, ACCUM_FP32, THERMOSTAT_NONE, RunParams
    from quantom_amd.engine import Engine

    def rel_rms(a, b):
        return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))

    out = {"mode": "accuracy"}
    prod = systems.production()
    p = RunParams(nlist_interval=0)
    o64 = Oracle(prod, p, precision=64, fast=True)
    o32 = Oracle(prod, p, precision=32, fast=True)
    for k in (KIND_CC, KIND_CS, KIND_SC, KIND_SS):
        o64.set_neighbors(k, o32.neighbors(k))
    f64c, f64s, _ = o64.compute_forces()
    ref = np.concatenate([f64c, f64s])
    for name, acc in (("fp32", ACCUM_FP32), ("fixed64", ACCUM_FIXED64)):
        e = Engine([prod], RunParams(nlist_interval=0, accumulation=acc))
        e.set_neighbors(cc=o32.neighbors(KIND_CC), cs=o32.neighbors(KIND_CS), ss=o32.neighbors(KIND_SS))
        cf, sf = e.compute_forces()
        out[f"production_rms_force_error_{name}"] = rel_rms(np.concatenate([cf, sf]), ref)
        e.close()
    log(f"force errors: {out}")

    def slope(t_ns, s, n_part):
        return float(np.polyfit(t_ns, np.asarray(s) / n_part, 1)[0])

    # production energy drift: equilibrated with the thermostat on the GPU, then NVE from that
    # same state on the GPU (fp32, fixed64) and in the FP64 oracle (a shorter window)
    eq = Engine([prod], RunParams(nlist_interval=10))
    eq.step(args.drift_equilibrate)
    cx, cxm, sx, sxm = eq.get_state()
    eq.close()
    ps = prod.copy()
    ps.chain_positions, ps.chain_prev_positions = cx.copy(), cxm.copy()
    ps.solvent_positions, ps.solvent_prev_positions = sx.copy(), sxm.copy()
    n_part = ps.n_chain + ps.n_solvents
    every = 100
    prod_series = {}
    for name, acc in (("fp32", ACCUM_FP32), ("fixed64", ACCUM_FIXED64)):
        e = Engine([ps], RunParams(nlist_interval=10, thermostat=THERMOSTAT_NONE, accumulation=acc))
        recs = e.step_observed(args.drift_steps_production, every)
        prod_series[name] = [r[0]["potential"] + r[0]["kinetic"] for r in recs]
        e.close()
    o_every = 50
    o = Oracle(ps, RunParams(nlist_interval=10, thermostat=THERMOSTAT_NONE), precision=64, fast=True)
    es = []
    for k in range(0, args.drift_steps_oracle, o_every):
        en = o.energies()
        es.append(en["potential"] + en["kinetic"])
        o.step(o_every)
    t_o = np.arange(len(es)) * o_every * ps.dt
    out["production_energy_drift_fp64_oracle_J_per_mol_per_particle_per_ns"] = slope(t_o, es, n_part)
    for name, ser in prod_series.items():
        t_g = np.arange(len(ser)) * every * ps.dt
        out[f"production_energy_drift_{name}_J_per_mol_per_particle_per_ns"] = slope(t_g, ser, n_part)
        m = int(np.searchsorted(t_g, t_o[-1], side="right"))
        out[f"production_energy_drift_{name}_oracle_window_J_per_mol_per_particle_per_ns"] = slope(t_g[:m], ser[:m], n_part)
    out["production_drift"] = {"equilibration_steps": args.drift_equilibrate, "gpu_steps": args.drift_steps_production,
                               "oracle_steps": args.drift_steps_oracle, "record_every_gpu": every,
                               "record_every_oracle": o_every, "mean_energy_J_per_mol_per_particle":
                               float(np.mean(prod_series["fp32"]) / n_part)}
    log(f"production drift: {out}")

    # energy drift: NVE, small system, GPU (fp32, fixed64) and the FP64 oracle
    small = systems.small()
    steps, every = args.drift_steps, 50
    pn = RunParams(nlist_interval=10, thermostat=THERMOSTAT_NONE)
    ns = small.dt * 1e-0
    series = {}
    for name, acc in (("fp32", ACCUM_FP32), ("fixed64", ACCUM_FIXED64)):
        e = Engine([small], RunParams(nlist_interval=10, thermostat=THERMOSTAT_NONE, accumulation=acc))
        es = []
        for k in range(0, steps + 1, every):
            en = e.energies()[0]
            es.append(en["potential"] + en["kinetic"])
            if k < steps:
                e.step(every)
        series[name] = es
        e.close()
    o = Oracle(small, pn, precision=64, fast=True)
    es = []
    for k in range(0, steps + 1, every):
        en = o.energies()
        es.append(en["potential"] + en["kinetic"])
        if k < steps:
            o.step(every)
    series["fp64_oracle"] = es
    t_ns = np.arange(len(es)) * every * small.dt
    n_part = small.n_chain + small.n_solvents
    for name, ser in series.items():
        out[f"small_energy_drift_{name}_J_per_mol_per_particle_per_ns"] = slope(t_ns, ser, n_part)
    out["drift_window_steps"] = steps
    out["note"] = ("angle and dihedral forces of the reference are not gradients of their energies, "
                   "so the drift is compared with the FP64 oracle's own drift")
    return out


def adjoint_bench(args):
    """64 replicas of 2,000 particles: gradients of an energy/force MSE over 200 frames (checkpoints every
    20), targets from the same engine at seeded perturbed parameters (sigma +-3%, epsilon +-10%).
    Reports gradients per second, the cost of the same gradients by central differences of forward runs,
    and a check of replica 0 against FP64 central differences of the oracle."""
    import numpy as np

    from quantom_amd import systems
    from quantom_amd.abi import THERMOSTAT_NONE, RunParams
    from quantom_amd.engine import Engine

    K, R, cpi = args.adj_frames, args.adj_replicas, args.adj_checkpoint
    syss = systems.adjoint_replicas(R)
    params = RunParams(nlist_interval=10, thermostat=THERMOSTAT_NONE)
    eng = Engine(syss, params)
    base = np.stack([eng.get_parameters(r) for r in range(R)])
    n_types = 16
    lj = [t for t in range(1, n_types + 1)] + [32 + t for t in range(1, n_types + 1)]   # sigma, epsilon of the 16 chain types
    theta = lj + [64, 65, 66, 67]
    rng = np.random.default_rng(0xAD1)
    pert = base.copy()
    for r in range(R):
        pert[r, 1:n_types + 1] *= 1.0 + 0.03 * rng.uniform(-1, 1, n_types)
        pert[r, 33:33 + n_types] *= 1.0 + 0.10 * rng.uniform(-1, 1, n_types)
        eng.set_parameters(pert[r], r)
    t0 = time.perf_counter()
    fc, fs, pot = eng.record_frames(K)
    log(f"targets: {K} frames of {R} replicas in {time.perf_counter() - t0:.1f}s")
    for r in range(R):
        eng.set_parameters(base[r], r)
    n_r = np.array([s.n_chain + s.n_solvents for s in syss])
    f2 = np.zeros(R)
    cb = np.cumsum([0] + [s.n_chain for s in syss])
    sb = np.cumsum([0] + [s.n_solvents for s in syss])
    for r in range(R):
        f2[r] = (fc[:, cb[r]:cb[r + 1]] ** 2).sum() + (fs[:, sb[r]:sb[r + 1]] ** 2).sum()
    beta = 1.0 / f2                                         # mean-squared force error, normalised
    alpha = 1.0 / (K * np.maximum(pot.var(axis=0), 1.0))    # mean-squared energy error, normalised
    eng.adjoint(K, alpha, beta, pot, fc, fs, checkpoint_interval=cpi)          # warm
    t0 = time.perf_counter()
    loss, grad = eng.adjoint(K, alpha, beta, pot, fc, fs, checkpoint_interval=cpi)
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    eng.adjoint(K, alpha, beta, pot, fc, fs, checkpoint_interval=cpi, gradient=False)
    t_fwd = time.perf_counter() - t0
    n_fd = 2 * len(theta)
    out = {"metric": "adjoint grads/sec", "value": R / t_adj, "unit": "gradients/s (36 parameters each)",
           "higher_is_better": True, "n_gpus": 1, "dtype": "f32 (f64 reductions)", "data": "synthetic",
           "config": {"workload": f"adjoint: {R} replicas x (512 chain + 1488 solvent), 4.2 nm, 16 types",
                      "frames": K, "checkpoint_interval": cpi, "nlist_interval": 10, "parameters": len(theta)},
           "adjoint_s_per_call": t_adj, "forward_loss_s_per_call": t_fwd,
           "fd_equivalent_grads_per_s": R / (n_fd * t_fwd), "adjoint_over_fd": n_fd * t_fwd / t_adj,
           "mean_loss": float(loss.mean())}
    log(f"adjoint {t_adj:.3f}s/call ({R / t_adj:.1f} grads/s), forward {t_fwd:.3f}s; FD would take {n_fd} forward runs")
    if args.adj_fd > 0:
        from oracle.oracle import Oracle
        s0 = syss[0]
        sel = [1, 2, 3, 33, 34, 35, 0, 32, 64, 65, 66, 67][:args.adj_fd]

        def oloss(v):
            s = s0.copy()
            s.forcefield = s.forcefield.copy()
            s.forcefield[:, 1] = v[:32]
            s.forcefield[:, 2] = v[32:64]
            for a in (s.bonds, s.bridge_bonds):
                a["kb"] = v[64]
                a["b0"] = v[65]
            for a in (s.angles, s.bridge_angles):
                a["k_theta"] = v[66]
                a["theta_0"] = v[67]
            o = Oracle(s, params, precision=64, fast=True)
            L = 0.0
            for k in range(K):
                c, so, _ = o.compute_forces()
                e = o.energies()["potential"]
                L += alpha[0] * (e - pot[k, 0]) ** 2 + beta[0] * (((c - fc[k, :cb[1]]) ** 2).sum() +
                                                                   ((so - fs[k, :sb[1]]) ** 2).sum())
                if k + 1 < K:
                    o.step(1)
            return L

        t0 = time.perf_counter()
        rows = []
        for i in sel:
            vp, vm = base[0].copy(), base[0].copy()
            h = 1e-5 * abs(base[0, i])
            vp[i] = np.float32(base[0, i] + h)
            vm[i] = np.float32(base[0, i] - h)
            fd = (oloss(vp) - oloss(vm)) / (float(vp[i]) - float(vm[i]))
            rows.append({"param": int(i), "adjoint": float(grad[0, i]), "fd_fp64": fd,
                         "rel_err": abs(grad[0, i] - fd) / max(abs(fd), 1e-300)})
            log(f"  param {i}: adjoint {grad[0, i]:.6e} fd {fd:.6e}")
        out["fd_check_replica0"] = rows
        out["fd_check_max_rel_err"] = max(r["rel_err"] for r in rows)
        out["fd_check_seconds"] = time.perf_counter() - t0
    return out


def copy_probe(args):
    """A plain device copy of known size (the scale for rocprofv3 FETCH_SIZE / WRITE_SIZE readings)."""
    import torch
    n = args.copy_mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.copy_reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"mode": "copy", "bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4,
            "launches": args.copy_reps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", default="bench", choices=["bench", "accuracy", "copy", "adjoint"])
    ap.add_argument("--adj-frames", type=int, default=200)
    ap.add_argument("--adj-replicas", type=int, default=64)
    ap.add_argument("--adj-checkpoint", type=int, default=0,
                    help="frames per checkpointed segment (0: one segment, nothing recomputed, when it fits)")
    ap.add_argument("--adj-fd", type=int, default=12, help="parameters of replica 0 checked by FP64 differences")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--copy-reps", type=int, default=10)
    ap.add_argument("--drift-steps", type=int, default=1000)
    ap.add_argument("--drift-steps-production", type=int, default=10000)
    ap.add_argument("--drift-steps-oracle", type=int, default=500)
    ap.add_argument("--drift-equilibrate", type=int, default=2000)
    ap.add_argument("--sequential", action="store_true",
                    help="replicas workload: also time the replicas as sequential single-system engines")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape instead of timing the candidates (about 20 ms)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="production")
    ap.add_argument("--accumulation", default=None, choices=["fp32", "fixed64"],
                    help="default: fixed64 for the large workload, fp32 otherwise")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-demo", action="store_true", help="skip the timing of the reference's demo system")
    ap.add_argument("--kernel-reps", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported by this build "
              f"(multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)

    from quantom_amd.abi import ACCUM_FIXED64, ACCUM_FP32, RunParams
    from quantom_amd.engine import Engine

    if args.mode == "accuracy":
        print(json.dumps(accuracy_report(args)), flush=True)
        return
    if args.mode == "adjoint":
        print(json.dumps(adjoint_bench(args)), flush=True)
        return
    if args.mode == "copy":
        print(json.dumps(copy_probe(args)), flush=True)
        return

    t_gen = time.perf_counter()
    syss, desc = workload(args.workload)
    log(f"generated {args.workload} in {time.perf_counter() - t_gen:.1f}s")
    n_chain = sum(s.n_chain for s in syss)
    n_solv = sum(s.n_solvents for s in syss)
    if args.accumulation is None:
        args.accumulation = "fixed64" if args.workload == "large" else "fp32"
    params = run_params(args.workload, args.accumulation)
    eng = Engine(syss, params)
    if not args.no_autotune:
        t_at = time.perf_counter()
        eng.autotune(10)
        log(f"autotune {time.perf_counter() - t_at:.1f}s")
    log(f"engine ready: chain {n_chain} solvent {n_solv} config {eng.config()}")
    # the in-step clock of the all-pairs kernel is switched on before the warmup: switching it
    # re-captures the launch graphs, and the first stepping call captures them (the whole rebuild
    # cycle and one graph per kind of step), so that no capture falls inside the timed steps
    eng.set_kernel_clock(True)
    t_w = time.perf_counter()
    eng.step(1)
    first_call_ms = 1000.0 * (time.perf_counter() - t_w)
    if args.warmup > 1:
        eng.step(args.warmup - 1)
    # untimed steps up to the next rebuild-cycle boundary (state parity 0, a fresh list period),
    # so that the timed steps replay whole captured cycles as a long run does, then whole cycles
    # until the GPU has stepped for RAMP_S: its clocks rise over the first ~25 ms of continuous
    # work (20-step windows measured 0.280 -> 0.244 ms/step on production), so a short timed
    # region right after a short warmup would time the ramp.  Both counts are reported.
    iv = max(1, params.nlist_interval)
    cycle = iv if iv % 2 == 0 else 2 * iv
    align = (-eng.steps) % cycle
    if align:
        eng.step(align)
    ramp = 0
    t_r = time.perf_counter()
    while time.perf_counter() - t_r < RAMP_S:
        eng.step(cycle)
        ramp += cycle
    eng.sync()
    clk0 = eng.kernel_clock()
    t0 = time.perf_counter()
    eng.step(args.steps)          # synchronises at the end
    t1 = time.perf_counter()
    clk1 = eng.kernel_clock()
    sc_live_n = clk1[1] - clk0[1]
    sc_live_us = (clk1[0] - clk0[0]) / max(sc_live_n, 1)
    ms = 1000.0 * (t1 - t0) / args.steps
    dt_fs = syss[0].dt * 1e6
    ns_day = dt_fs * 86.4 / ms
    log(f"timed {args.steps} steps: {ms:.4f} ms/step, {ns_day:.1f} ns/day")

    # The reference's calling pattern, deviceMaster() then hostMaster() once per step
    # (Environment.cpp:105-116), as include/lima_engine.hpp maps it: one queued step per call
    # (a graph replay), one wait per list period.  Beside it, a wait after every step.
    n_pc = max(args.steps, 2 * iv)
    t_pc = time.perf_counter()
    for i in range(n_pc):
        eng.step_async(1)
        if (i + 1) % iv == 0:
            eng.sync()
    eng.sync()
    per_call_ms = 1000.0 * (time.perf_counter() - t_pc) / n_pc
    t_pc = time.perf_counter()
    for i in range(n_pc):
        eng.step(1)
    per_call_sync_ms = 1000.0 * (time.perf_counter() - t_pc) / n_pc
    log(f"per-step calls: {per_call_ms:.4f} ms/step queued, {per_call_sync_ms:.4f} ms/step waiting on each")

    kt = eng.time_kernels(args.kernel_reps)
    log(f"kernel times (us): {kt}")
    log(f"solvent_chain inside the timed steps: {sc_live_us:.2f} us over {sc_live_n} launches "
        f"(alone: {kt['solvent_chain']:.2f} us)")
    pairs_sc = sum(s.n_chain * s.n_solvents for s in syss)
    flops = FLOPS_PER_PAIR * pairs_sc
    # the launch duration inside the timed steps, where the kernel shares the GPU with the step's
    # other streams (device clock per launch); the same kernel launched alone is reported beside it
    t_sc = sc_live_us * 1e-6
    achieved = flops / t_sc / 1e12 if t_sc > 0 else 0.0
    t_alone = kt["solvent_chain"] * 1e-6
    # VALU-bound (pairwise LJ is not a contraction, so MFMA does not apply).  The headline fraction
    # is against the unpacked FP32 FMA rate this build can issue (packed FP32 is disabled); the
    # fraction of the published packed peak stands beside it.
    peak_unpacked = PEAK_FP32_TFLOPS / 2
    roof = {"bound": "valu", "compute_unit": "VALU FP32, unpacked FMA (no MFMA: pairwise LJ is not a contraction)",
            "kernel": "solvent_chain_kernel", "achieved": achieved, "peak": peak_unpacked, "unit": "TFLOP/s",
            "frac": achieved / peak_unpacked, "traffic": read_traffic(f"solvent_chain_kernel_{args.workload}"),
            "flops_per_launch": flops, "avg_launch_us": sc_live_us, "launches_timed": sc_live_n,
            "timing": "device clock per launch over the timed steps (lima_get_kernel_clock)",
            "packed_peak": PEAK_FP32_TFLOPS, "packed_peak_frac": achieved / PEAK_FP32_TFLOPS,
            "alone_avg_launch_us": kt["solvent_chain"],
            "alone_achieved": flops / t_alone / 1e12 if t_alone > 0 else 0.0}

    out = {
        "metric": "ms/step & ns/day, 1 MI355X (single + 256-replica batch); adjoint grads/sec",
        "value": ns_day, "unit": "ns/day", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.accumulation == "fp32" else "f32+i64acc",
        "data": "reference demo input, ingested" if args.workload == "t4lysozyme" else "synthetic",
        "config": {"workload": desc, "n_chain": n_chain, "n_solvents": n_solv, "n_systems": len(syss),
                   "dt_fs": dt_fs, "cutoff_nm": params.cutoff, "nlist_interval": params.nlist_interval,
                   "thermostat": {0: "none", 1: "berendsen", 2: "reference"}[params.thermostat],
                   "accumulation": args.accumulation, "launch": eng.config()},
        "roofline": roof,
        "kernels_us": kt,
        "first_call_ms": first_call_ms,
        "warmup_align_steps": align,
        "warmup_ramp_steps": ramp,
        "warmup_note": (f"untimed: the {args.warmup} warmup steps, {align} steps to a rebuild-cycle boundary, "
                        f"then {ramp} steps (whole cycles, {RAMP_S * 1000:.0f} ms) while the GPU clocks rise"),
        "drop_in_calls": {"steps": n_pc, "ms_per_step_queued": per_call_ms, "ms_per_step_waiting": per_call_sync_ms,
                          "pattern": "one lima_step_async(1) per step, lima_sync every list period "
                                     "(lima::Engine::deviceMaster/hostMaster); beside it lima_step(1) per step"},
    }
    if len(syss) > 1:
        out["replica_ns_per_day_each"] = ns_day
        out["value"] = ns_day * len(syss)          # whole job: simulated ns of all replicas per day
        out["unit"] = "replica-ns/day"
        if args.sequential:
            eng.close()
            ms_seq = sequential_baseline(syss, params, max(10, args.steps // 10))
            out["sequential_ms_per_step"] = ms_seq
            out["batched_over_sequential"] = ms_seq / ms
            log(f"sequential single-system engines: {ms_seq:.3f} ms per step of all replicas")
    if args.workload == "production" and not args.no_demo:
        eng.close()
        out["published_comparison"] = demo_comparison(2000, 200)
        log(f"demo: {out['published_comparison']['ms_per_step']:.4f} ms/step "
            f"(published {PUBLISHED_MS_PER_STEP} ms/step)")
    if not args.no_cpu and len(syss) == 1:
        cb = cpu_baseline(syss, params, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = cb["ms_per_step"] / ms
        log(f"cpu baseline: {cb['ms_per_step']:.1f} ms/step on {cb['cores']} threads")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Adjoint accuracy against the horizon: one replica of the adjoint workload (512 chain + 1,488
solvent, 4.2 nm, 16 types), a loss over per-frame potential energy and forces against targets at
perturbed parameters (sigma +-3%, epsilon +-10%), at 20, 50, 100 and 200 frames (checkpoints
every 20).  For eight parameters (two types' sigma and epsilon, kb, b0, k_theta, theta_0) it
prints the FP64 oracle's central differences at steps of 1e-2 .. 1e-6 relative (their spread
shows whether the differences have converged) and the GPU adjoint's gradient (FP32 and int64
accumulation) against the 1e-5 step.  Run from the repository root on a GPU:

    python tests/adjoint_horizon.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from quantom_amd import systems  # noqa: E402
from quantom_amd.abi import RunParams  # noqa: E402
from quantom_amd.engine import Engine  # noqa: E402
from tests.test_adjoint import PARAMS, base_params, oracle_frames, oracle_loss, used_types, with_params  # noqa: E402

STEPS = (1e-2, 1e-3, 1e-4, 1e-5, 1e-6)


def main():
    s = systems.adjoint_replicas(1)[0]
    v0 = base_params(s)
    types = used_types(s)
    rng = np.random.default_rng(0x51A7 + 4000)
    vt = v0.copy()
    for t in types:
        vt[t] *= 1.0 + 0.03 * rng.uniform(-1, 1)
        vt[32 + t] *= 1.0 + 0.10 * rng.uniform(-1, 1)
    params = [types[1], types[5], 32 + types[1], 32 + types[5], 64, 65, 66, 67]
    print("| frames | param | adjoint fp32 | adjoint int64 | FD 1e-5 | rel fp32 | rel int64 | FD spread 1e-3..1e-6 |")
    print("|---:|---:|---:|---:|---:|---:|---:|---:|")
    for K in (20, 50, 100, 200):
        fct, fst, et = oracle_frames(with_params(s, vt.astype(np.float32).astype(np.float64)), K)
        bfc, bfs, bpot = oracle_frames(s, K)
        beta = 1.0 / float((fct ** 2).sum() + (fst ** 2).sum())
        alpha = beta * float(((bfc - fct) ** 2).sum() + ((bfs - fst) ** 2).sum()) / float(((bpot - et) ** 2).sum())
        fct32, fst32 = fct.astype(np.float32), fst.astype(np.float32)
        grads = []
        for acc in (0, 1):
            e = Engine([s], RunParams(nlist_interval=PARAMS.nlist_interval, thermostat=PARAMS.thermostat,
                                      accumulation=acc))
            _, g = e.adjoint(K, [alpha], [beta], et[:, None], fct32, fst32, checkpoint_interval=20)
            e.close()
            grads.append(g[0])
        for i in params:
            fd = {}
            for rel in STEPS:
                h = rel * max(abs(v0[i]), 1e-3)
                vp, vm = v0.copy(), v0.copy()
                vp[i] = np.float32(v0[i] + h)
                vm[i] = np.float32(v0[i] - h)
                lp = oracle_loss(with_params(s, vp), K, alpha, beta, et, fct32, fst32)
                lm = oracle_loss(with_params(s, vm), K, alpha, beta, et, fct32, fst32)
                fd[rel] = (lp - lm) / (float(vp[i]) - float(vm[i]))
            ref = fd[1e-5]
            spread = max(abs(fd[r] - ref) for r in (1e-3, 1e-4, 1e-6)) / abs(ref)
            print(f"| {K} | {i} | {grads[0][i]:.6e} | {grads[1][i]:.6e} | {ref:.6e} | "
                  f"{abs(grads[0][i] - ref) / abs(ref):.1e} | {abs(grads[1][i] - ref) / abs(ref):.1e} | {spread:.1e} |",
                  flush=True)


if __name__ == "__main__":
    main()

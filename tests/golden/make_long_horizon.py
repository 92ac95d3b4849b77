"""Writes the long-horizon fixtures: how far float32 rollouts of the oracle (a
numpy restatement of jax-sph's si_euler + WCSPH.forward + neighbour list)
stay with float64 rollouts of the same algorithm, and float32 oracle states at
checkpoints, for two cases:

* "c1": 2D Taylor-Green vortex, dx = 0.02 (2,500 particles), standard SPH,
  tvf = 0, dt = 4e-4, 1000 steps;
* "tgv3d": 3D Taylor-Green vortex, 16^3 = 4,096 particles, tvf = 1, CFL dt,
  600 steps.

tests/golden/long_horizon.json holds, per case, the field-relative error
max|a32 - a64| / max|a64| of r, u, v and rho every 10 steps, and K: the last
step up to which every recorded error is at most 2e-4 (after K the float32
and float64 trajectories part exponentially).  tests/golden/long_horizon_<case>.npz
holds the float32 oracle's r, u, v, rho at steps K/4, K/2, 3K/4 and K, which
the GPU tests compare against (1e-3 relative).

Run from the repository root: ``python tests/golden/make_long_horizon.py``
(about 30 CPU-minutes; the four rollouts run in parallel processes).
"""
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "jax-sph_amd"))

CASES = {
    "c1": ("tgv", {"case": {"dim": 2, "dx": 0.02}, "solver": {"dt": 0.0004}}, 1000),
    "tgv3d": ("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / 16}, "solver": {"tvf": 1.0}}, 600),
}
EVERY = 10
FIELDS = ("r", "u", "v", "rho")
AGREE = 2e-4


def rollout(job):
    name, f64 = job
    from jax_sph_amd import case_setup as cs
    from oracle import nl_oracle as nlo
    from oracle import sph_oracle as so
    from tests import gpu_helpers as G
    case, ov, steps = CASES[name]
    cfg = cs.load_config(case, ov)
    (cfg, box, state, g, bc, nw, eos, key, disp, shift) = cs.load_case(case)(cfg).initialize("cpu")
    solver = cs.make_solver(cfg, disp, eos, g)
    bcf = G.oracle_bc(bc, cfg.case.dim)
    st = G.to_np(state)
    if f64:
        ns = nlo.NeighborSearch(box, solver._kernel_fn.cutoff, mask_self=False, dtype=np.float64)
        adv = so.si_euler(cfg.solver.tvf, G.oracle_solver64(solver, box).forward, box, bcf,
                          ns.update, np.float64)
        st = G.to_f64(st)
    else:
        ns = nlo.NeighborSearch(box, solver._kernel_fn.cutoff, mask_self=False)
        adv = so.si_euler(cfg.solver.tvf, G.oracle_solver(solver, box).forward, box, bcf,
                          ns.update)
    nb = ns.allocate(st["r"])
    out = {}
    for k in range(1, steps + 1):
        st, nb = adv(cfg.solver.dt, st, nb)
        if k % EVERY == 0:
            out[k] = {f: st[f].copy() for f in FIELDS}
    return name, f64, out, float(cfg.solver.dt)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


if __name__ == "__main__":
    t0 = time.time()
    jobs = [(n, f) for n in CASES for f in (False, True)]
    with mp.Pool(len(jobs)) as pool:
        res = pool.map(rollout, jobs, chunksize=1)
    runs = {(n, f): (o, dt) for n, f, o, dt in res}
    summary = {}
    for name, (case, ov, steps) in CASES.items():
        o32, dt = runs[(name, False)]
        o64, _ = runs[(name, True)]
        curve = {f: [rel(o32[k][f], o64[k][f]) for k in sorted(o32)] for f in FIELDS}
        ks = sorted(o32)
        K = 0
        for i, k in enumerate(ks):
            if max(curve[f][i] for f in FIELDS) > AGREE:
                break
            K = k
        checkpoints = sorted({max(EVERY, (K * q // 4) // EVERY * EVERY) for q in (1, 2, 3, 4)})
        np.savez(os.path.join(ROOT, "tests", "golden", f"long_horizon_{name}.npz"),
                 steps=np.array(checkpoints),
                 **{f"{f}_{k}": o32[k][f].astype(np.float32) for k in checkpoints for f in FIELDS})
        summary[name] = {"case": case, "overrides": ov, "dt": dt, "steps": steps,
                         "every": EVERY, "agree": AGREE, "K": K, "checkpoints": checkpoints,
                         "steps_recorded": ks, "rel_err_f32_vs_f64": curve}
        print(name, "K =", K, {f: curve[f][-1] for f in FIELDS}, flush=True)
    summary["seconds"] = time.time() - t0
    json.dump(summary, open(os.path.join(ROOT, "tests", "golden", "long_horizon.json"), "w"))

"""The 2D Taylor-Green vortex of the CPU-runnable configuration (dx = 0.02,
2,500 particles, standard SPH, no background pressure, dt = 4e-4) is chaotic
over its 1000 steps: float32 and float64 rollouts of the same algorithm (the
oracle) agree to 2e-4 up to step K (tests/golden/long_horizon.json) and then
part exponentially, by more than 1e-3 before step 400.  This is why the GPU
rollout is held to the float32 oracle only up to K
(tests/test_long_horizon_gpu.py), with kinetic energy and finiteness checked
beyond it."""
import json
import os

import numpy as np

from tests import gpu_helpers as G

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_c1_float32_and_float64_oracles_part_after_K():
    from jax_sph_amd import case_setup as cs
    from oracle import nl_oracle as nlo
    from oracle import sph_oracle as so
    fx = json.load(open(os.path.join(GOLDEN, "long_horizon.json")))["c1"]
    cfg = cs.load_config(fx["case"], fx["overrides"])
    (cfg, box, state, g, bc, nw, eos, key, disp, shift) = \
        cs.load_case(fx["case"])(cfg).initialize("cpu")
    solver = cs.make_solver(cfg, disp, eos, g)
    bcf = G.oracle_bc(bc, 2)
    ns = nlo.NeighborSearch(box, solver._kernel_fn.cutoff, mask_self=False)
    ns64 = nlo.NeighborSearch(box, solver._kernel_fn.cutoff, mask_self=False, dtype=np.float64)
    adv = so.si_euler(0.0, G.oracle_solver(solver, box).forward, box, bcf, ns.update)
    adv64 = so.si_euler(0.0, G.oracle_solver64(solver, box).forward, box, bcf, ns64.update,
                        np.float64)
    st = G.to_np(state)
    st64 = G.to_f64(st)
    nb, nb64 = ns.allocate(st["r"]), ns64.allocate(st64["r"])
    K = fx["K"]
    errs = {}
    for k in range(1, 401):
        st, nb = adv(cfg.solver.dt, st, nb)
        st64, nb64 = adv64(cfg.solver.dt, st64, nb64)
        if k % 10 == 0:
            errs[k] = max(G.rel_err(st[f], st64[f]) for f in ("u", "v", "rho"))
    assert max(e for k, e in errs.items() if k <= K) <= 2e-4, errs
    assert errs[400] > 1e-3, errs
    # u is the field that parts first; the committed curve says the same
    assert fx["rel_err_f32_vs_f64"]["u"][39] > 1e-3

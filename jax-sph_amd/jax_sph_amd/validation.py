"""Validation runs of the fused step (validation/validate.py of the reference).

The reference's validation scripts run long simulations with ``main.py``,
write h5 trajectories and plot them against published data
(validation/tgv3d.sh, ldc2d.sh, validate.py).  Here the same quantities are
computed in memory, without files or plots:

* :func:`tgv_energy`: kinetic energy per volume E(t) = get_ekin / L^d and
  -dE/dt between records of a Taylor-Green vortex (validate.py:107-138,
  compared there with JAX-Fluids data, validation/ref/tgv3d_ref_*.txt);
* :func:`ldc_centrelines`: the lid-driven cavity's u(y) and v(x) along the
  centrelines through ``sph_interpolator`` at the end of the run
  (validate.py:318-356, compared there with Ghia et al. 1982).

Like the reference's scripts, each run starts from relaxed positions
(:func:`relaxed_state`: the case at rest from noisy positions, 5000 steps).

Steps run as HIP-graph replays of ``every`` fused steps.  A list overflow
inside a replay re-allocates the list from the state at the start of that
replay and redoes it, as simulate.py:119-130 does per step.
"""
from __future__ import annotations

from typing import Callable, Dict, Tuple

import numpy as np
import torch

from .case_setup import Cfg, load_case, load_config, make_solver
from .graph import StepGraph
from .integrator import si_euler
from .jax_md.partition import Sparse
from . import partition as _partition
from .simulate import simulate
from .utils import get_ekin, sph_interpolator


def relaxed_state(case: str, overrides: Dict, p_bg_factor: float) -> Dict:
    """The relaxation the reference's validation scripts run first
    (validation/tgv2d.sh:13, tgv3d.sh:18, ldc2d.sh:7): the case at rest from
    positions with 0.25 dx noise, transport velocity on, seed 123, background
    pressure factor ``p_bg_factor``, 5000 steps.  Returns the final state."""
    cfg = load_config(case, overrides)
    cfg.case.mode = "rlx"
    cfg.case.r0_noise_factor = 0.25
    cfg.solver.tvf = 1.0
    cfg.solver.name = "SPH"
    cfg.solver.dt = None if case == "tgv" else cfg.solver.dt
    cfg.eos.p_bg_factor = p_bg_factor
    cfg.seed = 123
    return simulate(cfg, verbose=False)["state"]


def run(cfg: Cfg, every: int, record: Callable[[int, Dict], None],
        state0: Dict = None) -> Tuple[Cfg, Dict]:
    """Runs ``cfg`` for its ``t_end`` (rounded down to whole records of
    ``every`` steps), calling ``record(step, state)`` at step 0 and after every
    record; ``state0`` supplies relaxed positions.  Returns the set-up config
    (with dt) and the final state."""
    Case = load_case(cfg.case.source or cfg.config)
    (cfg, box, state, g_ext_fn, bc_fn, nw_fn, eos, key, disp, shift) = Case(cfg).initialize(
        state0=state0)
    solver = make_solver(cfg, disp, eos, g_ext_fn)
    fn = _partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                  capacity_multiplier=1.25, mask_self=False, format=Sparse)
    advance = si_euler(cfg.solver.tvf, solver.forward_wrapper(), shift, bc_fn, nw_fn)
    dt = cfg.solver.dt
    n_records = int(cfg.solver.t_end / dt) // every
    record(0, state)
    nb = fn.allocate(state["r"])
    graph = StepGraph(advance, state, nb, dt, steps_per_replay=every)
    for k in range(1, n_records + 1):
        start = {key_: v.clone() for key_, v in graph.state.items()}
        graph.replay()
        if graph.did_buffer_overflow:
            # redo the chunk one step at a time, re-allocating where a step
            # overflows (simulate.py:119-130), then capture again from there
            st, nb = start, graph.neighbors
            nb = fn.allocate(st["r"])
            for _ in range(every):
                st_new, nb_new = advance(dt, st, nb)
                if bool(nb_new.did_buffer_overflow):
                    nb = fn.allocate(st["r"])
                    st_new, nb_new = advance(dt, st, nb)
                    if bool(nb_new.did_buffer_overflow):
                        raise RuntimeError(f"neighbour list overflow after re-allocation "
                                           f"in record {k}")
                st, nb = st_new, nb_new
            graph = StepGraph(advance, st, nb, dt, steps_per_replay=every)
        record(k * every, graph.state)
    return cfg, graph.state


def tgv_energy(dim: int = 3, nx: int = 32, Re: float = 100.0, tvf: float = 1.0,
               t_end: float = 10.0, every: int = 50, solver: str = "SPH",
               relaxed: bool = True) -> Dict:
    """Taylor-Green vortex (from relaxed positions, or the lattice): times,
    E(t) = E_kin / L^d and -dE/dt (forward differences between records), as
    validate.py:107-138."""
    side = 2 * np.pi if dim == 3 else 1.0
    case = {"dim": dim, "dx": side / nx, "viscosity": 1.0 / Re}
    state0 = relaxed_state("tgv", {"case": dict(case)}, 0.02) if relaxed else None
    cfg = load_config("tgv", {"case": case,
                              "solver": {"name": solver, "tvf": tvf, "t_end": t_end}})
    ts, es = [], []

    def rec(step, state):
        ts.append(step)
        es.append(get_ekin(state, cfg.case.dx) / side ** dim)

    cfg, _ = run(cfg, every, rec, state0)
    t = np.asarray(ts, dtype=np.float64) * cfg.solver.dt
    e = np.asarray(es)
    return {"t": t, "E": e, "minus_dEdt": -(e[1:] - e[:-1]) / (t[1:] - t[:-1]),
            "dt": cfg.solver.dt, "dx": cfg.case.dx}


def ldc_centrelines(dx: float = 0.02, t_end: float = 20.0, solver: str = "SPH",
                    tvf: float = 0.0, every: int = 500, relaxed: bool = True,
                    average: float = 5.0) -> Dict:
    """Lid-driven cavity at Re = 100 (ldc.yaml, dt = 4e-4): u along the
    vertical and v along the horizontal centreline, interpolated as
    validate.py:318-356 (points dx apart from the fluid box's corner, offset by
    the 3 wall layers).

    The reference plots the last state.  The steady cavity carries particle
    noise: at dx = 0.02 single snapshots between t = 15 and 20 differ by
    0.02-0.05 on the centrelines.  The profiles returned here are therefore
    averaged over the snapshots at whole time units in the last ``average``
    time units (t_end - average ... t_end); ``average = 0`` gives the last
    state alone."""
    state0 = relaxed_state("ldc", {"case": {"dx": dx}}, 0.0) if relaxed else None
    cfg = load_config("ldc", {"case": {"dx": dx},
                              "solver": {"name": solver, "tvf": tvf, "t_end": t_end}})
    unit = int(round(1.0 / cfg.solver.dt))
    if unit % every:
        raise ValueError(f"every = {every} must divide the {unit} steps of one time unit")
    first = int(round((t_end - average) / cfg.solver.dt))
    snaps = []

    def rec(step, st):
        if step >= first and step % unit == 0:
            snaps.append({k: v.clone() for k, v in st.items()})

    cfg, state = run(cfg, every, rec, state0)
    if not snaps:
        snaps = [state]
    n = round(1 / dx) + 1
    axis = np.array([dx * (3 + i) for i in range(n)])
    us, vs = [], []
    for st in snaps:
        interp = sph_interpolator(cfg, st, "vector")
        rs = (0.5 + 3 * dx) * np.ones((n, 2))
        rs[:, 1] = axis
        us.append(interp(st, rs, "u", dim_ind=0).cpu().numpy())
        rs = (0.5 + 3 * dx) * np.ones((n, 2))
        rs[:, 0] = axis
        vs.append(interp(st, rs, "u", dim_ind=1).cpu().numpy())
    return {"y": axis - 3 * dx, "u": np.mean(us, axis=0), "x": axis - 3 * dx,
            "v": np.mean(vs, axis=0), "snapshots": len(snaps)}

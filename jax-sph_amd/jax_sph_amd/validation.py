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
  (validate.py:318-356, compared there with Ghia et al. 1982);
* :func:`rpf_profiles`: the reverse Poiseuille flow's velocity profile at
  given times (cases/rpf.py; the reference ships such profiles as
  validation/ref/rpf_t_*.csv), with :func:`rpf_series`, the flow's
  closed-form start-up solution;
* :func:`db_front`: the dam break's surge front over time (validation/db2d.sh's
  Riemann settings on a column twice as high as wide), the form of Martin &
  Moyce's (1952) data in validation/ref/dambreak_data.py.

Like the reference's scripts, the TGV and cavity runs start from relaxed positions
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


def rpf_series(y, t: float, g: float = 1.0, nu: float = 0.1, terms: int = 400) -> np.ndarray:
    """Start-up from rest of the flow driven by the square-wave force of
    cases/rpf.py (+g for 0 < y < 1, -g for 1 < y < 2, period 2 in y) in
    u_t = nu u_yy + f(y): with f = sum over odd k of 4g/(k pi) sin(k pi y),
    u(y, t) = sum over odd k of 4g/(k pi) (1 - exp(-nu (k pi)^2 t)) / (nu (k pi)^2)
    sin(k pi y).  As t grows it tends to g/(2 nu) y (1 - y) on [0, 1]."""
    y = np.asarray(y, np.float64)
    k = np.arange(1, 2 * terms, 2, dtype=np.float64)[:, None]
    lam = nu * (k * np.pi) ** 2
    amp = 4.0 * g / (k * np.pi) * (-np.expm1(-lam * t)) / lam
    return (amp * np.sin(k * np.pi * y[None, :])).sum(axis=0)


def rpf_profiles(times=(0.2, 1.0, 2.0, 10.0), dx: float = 0.025, solver: str = "SPH",
                 tvf: float = 0.0) -> Dict:
    """Reverse Poiseuille flow (rpf.yaml: dx = 0.025, dt = 5e-4, nu = 0.1,
    g = 1) from the lattice at rest, run to each of ``times``.  The flow stays
    parallel to x, so each lattice row keeps its y; the profile at a time is
    the mean u_x of each row, the upper half's rows (1 < y < 2) negated and
    laid over the lower half's (the force, and so the flow, is antisymmetric
    there).  Returns the rows' y in (0, 1), the profile per time, and the
    largest |u_y|.  The force depends on position, so the step runs as
    separate library calls rather than graph-replayed."""
    cfg = load_config("rpf", {"case": {"dx": dx},
                              "solver": {"name": solver, "tvf": tvf, "t_end": max(times)}})
    Case = load_case("rpf")
    (cfg, box, state, g_ext_fn, bc_fn, nw_fn, eos, key, disp, shift) = Case(cfg).initialize()
    model = make_solver(cfg, disp, eos, g_ext_fn)
    fn = _partition.neighbor_list(disp, box, r_cutoff=model._kernel_fn.cutoff,
                                  capacity_multiplier=1.25, mask_self=False, format=Sparse)
    advance = si_euler(cfg.solver.tvf, model.forward_wrapper(), shift, bc_fn, nw_fn)
    dt = cfg.solver.dt
    marks = {int(round(t / dt)): t for t in times}
    rows = torch.floor(state["r"][:, 1] / dx).long()
    n_half = int(round(1.0 / dx))
    if int(rows.max()) != 2 * n_half - 1:
        raise ValueError("the lattice rows do not fill the box")
    sign = torch.where(rows < n_half, 1.0, -1.0).to(state["u"].dtype)
    half_row = torch.remainder(rows, n_half)
    counts = torch.bincount(half_row, minlength=n_half).to(state["u"].dtype)
    out, uy_max = {}, 0.0
    st, nb = state, fn.allocate(state["r"])
    for step in range(1, max(marks) + 1):
        st_new, nb_new = advance(dt, st, nb)
        if bool(nb_new.did_buffer_overflow):          # simulate.py:119-130
            nb = fn.allocate(st["r"])
            st_new, nb_new = advance(dt, st, nb)
            if bool(nb_new.did_buffer_overflow):
                raise RuntimeError(f"neighbour list overflow after re-allocation at step {step}")
        st, nb = st_new, nb_new
        if step in marks:
            prof = torch.zeros(n_half, dtype=st["u"].dtype, device=st["u"].device)
            prof.index_add_(0, half_row, sign * st["u"][:, 0])
            out[marks[step]] = (prof / counts).cpu().numpy()
            uy_max = max(uy_max, float(st["u"][:, 1].abs().max()))
    y = (np.arange(n_half) + 0.5) * dx
    return {"y": y, "u": out, "uy_max": uy_max, "dt": dt, "nu": cfg.case.viscosity,
            "g": cfg.case.g_ext_magnitude}


def db_front(dx: float = 0.02, solver: str = "RIE", T_end: float = 3.4, every: int = 25,
             free_slip: bool = True, n2: float = 2.0) -> Dict:
    """Dam break of a column of width a = 1 and height ``n2`` a (the db case
    with special.L = 1, special.H = n2, in a 5.366-long tank at least n2 + 0.5
    high) with the settings of validation/db2d.sh's Riemann run: u_ref = 2, no
    viscosity, free-slip walls, dt = 2e-4 at dx = 0.02 (scaled with dx),
    limiter 3, no artificial viscosity.  Runs to T = t sqrt(2 g / a) =
    ``T_end`` and returns the surge front's distance from the left wall over
    a, Z(T) (the farthest fluid particle, less the wall layers), every
    ``every`` steps: the form of Martin & Moyce's (1952) data for their
    n^2 = 2 column."""
    cfg = load_config("db", {"case": {"dx": dx, "u_ref": 2.0, "viscosity": 0.0,
                                      "special": {"L": 1.0, "H": n2,
                                                  "H_wall": max(2.0, n2 + 0.5)}},
                             "solver": {"name": solver, "dt": 0.0002 * dx / 0.02,
                                        "free_slip": free_slip,
                                        "artificial_alpha": 0.0, "eta_limiter": 3}})
    g, a = cfg.case.g_ext_magnitude, 1.0
    cfg.solver.t_end = T_end / np.sqrt(2.0 * g / a)
    wall = cfg.solver.n_walls * dx
    ts, zs = [], []

    def rec(step, st):
        fluid = st["tag"] == 0
        ts.append(step)
        zs.append((float(st["r"][fluid, 0].max()) - wall) / a)

    cfg, _ = run(cfg, every, rec)
    T = np.asarray(ts, dtype=np.float64) * cfg.solver.dt * np.sqrt(2.0 * g / a)
    return {"T": T, "Z": np.asarray(zs), "dt": cfg.solver.dt, "dx": dx}

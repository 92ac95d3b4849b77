"""Host step loop (drop-in for ``jax_sph/simulate.py:19-136``).

Same order as the reference:
1. case set-up, solver, neighbour list and ``si_euler``;
2. one warm-up step whose result is discarded (``simulate.py:110-111``);
3. ``sequence_length + 2`` steps, each preceded by ``write_state(step - 1)``.

On a neighbour-list overflow, the loop re-allocates from the pre-step
positions and redoes the step (``simulate.py:119-130``), so trajectory entry
``k`` is the state after ``k + 1`` advances.  State writing is a callback
(``write_fn(step, state)``): the reference's HDF5 writer (``io_state.py``)
needs ``h5py``, which this package does not depend on.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, Optional

import numpy as np
import torch

from . import partition
from .case_setup import Cfg, load_case, make_solver
from .integrator import si_euler
from .jax_md.partition import Sparse
from .utils import Logger


def simulate(cfg: Cfg, write_fn: Optional[Callable[[int, Dict], None]] = None,
             verbose: bool = True, state0: Optional[Dict] = None) -> Dict:
    """Runs a case; returns the final state and run statistics.

    ``cfg.case.mode == "rlx"`` runs the case's relaxation (simulate.py:30-32,
    case_setup.py:356-404: 5000 steps at rest from noisy positions); its final
    state is the ``state0`` (positions) of a later run."""
    Case = load_case(cfg.case.source or cfg.config)
    (cfg, box, state, g_ext_fn, bc_fn, nw_fn, eos, key, displacement_fn,
     shift_fn) = Case(cfg).initialize(state0=state0)
    if hasattr(write_fn, "bind"):  # a writer that sets up its directory (io_state.py:14-27)
        write_fn.bind(cfg)
    solver = make_solver(cfg, displacement_fn, eos, g_ext_fn)
    forward = solver.forward_wrapper()
    n = state["r"].shape[0]
    neighbor_fn = partition.neighbor_list(
        displacement_fn, box, r_cutoff=solver._kernel_fn.cutoff, backend=cfg.nl.backend,
        capacity_multiplier=1.25, mask_self=False, format=Sparse, num_particles_max=n,
        num_partitions=cfg.nl.num_partitions, pbc=np.array(cfg.case.pbc))
    neighbors = neighbor_fn.allocate(state["r"])
    advance = si_euler(cfg.solver.tvf, forward, shift_fn, bc_fn, nw_fn)
    logger = Logger(dt=cfg.solver.dt, dx=cfg.case.dx, print_props=cfg.io.print_props,
                    sequence_length=cfg.solver.sequence_length)

    # warm-up: initialises the accelerations of the fused path; discarded
    _state, _neighbors = advance(0.0, state, neighbors)
    torch.cuda.synchronize()
    del _state, _neighbors

    reallocations = 0
    start = time.time()
    for step in range(cfg.solver.sequence_length + 2):
        if write_fn is not None:
            write_fn(step - 1, state)
        state_, neighbors_ = advance(cfg.solver.dt, state, neighbors)
        if bool(neighbors_.did_buffer_overflow):
            if verbose:
                print(f"Reallocate neighbors list {(2, neighbors.max_occupancy)} at step {step}")
            neighbors = neighbor_fn.allocate(state["r"])
            reallocations += 1
            if verbose:
                print(f"To list {(2, neighbors.max_occupancy)}")
            state, neighbors = advance(cfg.solver.dt, state, neighbors)
        else:
            state, neighbors = state_, neighbors_
        if verbose and step % cfg.io.write_every == 0:
            logger.print_stats(state, step)
    torch.cuda.synchronize()
    elapsed = time.time() - start
    if verbose:
        print(f"time: {elapsed:.2f} s")
    return {"state": state, "neighbors": neighbors, "reallocations": reallocations,
            "elapsed_s": elapsed, "cfg": cfg}

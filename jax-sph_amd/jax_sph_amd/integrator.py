"""Semi-implicit Euler with transport velocity (drop-in for ``jax_sph/integrator.py``).

``si_euler(tvf, model, shift_fn, bc_fn, nw_fn)`` returns
``advance(dt, state, neighbors) -> (state, neighbors)`` with the reference's
order (integrator.py:22-56): kick ``u += dt*dudt``; ``v = u + tvf*dt/2*dvdt``;
drift and wrap ``r = mod(r + dt*v, box)``; neighbour update; RHS; case BC.

When every piece is one this package understands -- ``model`` from
``WCSPH.forward_wrapper()``, ``shift_fn`` from ``space.periodic``, a
``CaseBC`` (or None) for ``bc_fn``, no moving-wall normals, a constant body
force and a neighbour list rebuilt every step (``dr_threshold == 0``) -- the
whole step is one ``sph_step`` call into the HIP library, on the list's
Verlet-skin buffers when it has them (same pair sets as a per-step rebuild).  Otherwise the same
HIP kernels run one call at a time and Python callables are applied between
them.  Either way the input dict is left untouched and a new dict is returned.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Dict

import numpy as np
import torch

from . import _lib
from .jax_md.partition import NeighborList
from .jax_md.space import PeriodicShift


_f = np.float32


def _bc_spec(bc_fn):
    from .case_setup import CaseBC
    spec = _lib.BCSpec()
    spec.kind = 0
    if bc_fn is None:
        return spec
    if isinstance(bc_fn, CaseBC):
        spec.kind = bc_fn.kind
        for d in range(3):
            spec.u_lid[d] = float(bc_fn.u_lid[d]) if d < len(bc_fn.u_lid) else 0.0
        return spec
    return None


def si_euler(tvf: float, model: Callable, shift_fn: Callable, bc_fn: Callable,
             nw_fn: Callable):
    """Semi-implicit Euler integrator including Transport Velocity."""

    solver = getattr(model, "_sph_solver", None)

    def _fusable(neighbors):
        if solver is None or nw_fn is not None or not isinstance(neighbors, NeighborList):
            return None
        if not isinstance(shift_fn, PeriodicShift) or not shift_fn.wrapped:
            return None
        if neighbors._search.dr_threshold != 0.0 or solver.g_const is None:
            return None
        if not np.array_equal(shift_fn.side, solver.displacement_fn.side):
            return None
        return _bc_spec(bc_fn)

    def _fused(dt, state, neighbors, spec):
        ws = neighbors.workspace
        neighbors._check_live()
        s = solver.state_in(state)
        if s["r"].shape[0] != ws.n:
            raise ValueError("state size does not match the neighbour list")
        out = solver.new_outputs(s)
        r_out = torch.empty_like(s["r"])
        bank = neighbors._search.step_bank(neighbors)
        ws.generation[bank] += 1
        info = torch.empty(8, dtype=torch.int64, device=r_out.device)
        dtf = _f(dt)
        ctvf = _f(_f(tvf * 0.5) * dtf)
        # v == u exactly when tvf == 0, so the A = rho u (v-u) stress vanishes
        prm = solver.params(tvf_stress=(tvf != 0.0))
        _lib.check(_lib.lib().sph_step(
            ws.handle, bank, ctypes.byref(prm), float(dtf), float(ctvf), ctypes.byref(spec),
            ctypes.byref(solver.c_state_in(s)), _lib.ptr(r_out),
            ctypes.byref(solver.c_state_out(out)), _lib.ptr(neighbors.info),
            neighbors.max_occupancy, _lib.ptr(info), _lib.stream()), "sph_step")
        nbrs = neighbors._search.adopt_step(neighbors, bank, info, r_out)
        s = dict(s)
        s["r"] = r_out
        return solver.assemble(s, out), nbrs

    def advance(dt: float, state: Dict, neighbors):
        """Call to integrator."""
        spec = _fusable(neighbors)
        if spec is not None:
            return _fused(dt, state, neighbors, spec)
        state = dict(state)
        dtf = _f(dt)
        ctvf = _f(_f(tvf * 0.5) * dtf)
        u = _lib.as_f32(state["u"])
        dudt = _lib.as_f32(state["dudt"])
        dvdt = _lib.as_f32(state["dvdt"])
        r = _lib.as_f32(state["r"])
        n, dim = r.shape
        u1, v1 = torch.empty_like(u), torch.empty_like(u)
        if isinstance(shift_fn, PeriodicShift) and shift_fn.wrapped:
            r1 = torch.empty_like(r)
            box = (ctypes.c_float * 3)(*([float(x) for x in shift_fn.side] + [0.0] * (3 - dim)))
            _lib.check(_lib.lib().sph_integrate(dim, n, _lib.ptr(r), _lib.ptr(u),
                                                _lib.ptr(dudt), _lib.ptr(dvdt), _lib.ptr(r1),
                                                _lib.ptr(u1), _lib.ptr(v1), float(dtf),
                                                float(ctvf), box, _lib.stream()),
                       "sph_integrate")
        else:
            u1 = u + float(dtf) * dudt
            v1 = u1 + float(ctvf) * dvdt
            r1 = shift_fn(r, float(dtf) * v1)
        state["u"], state["v"], state["r"] = u1, v1, r1
        if nw_fn is not None:
            state["nw"] = nw_fn(state["r"])
        # num_particles (integrator.py:305) only matters to the matscipy backend
        neighbors = neighbors.update(state["r"])
        state = model(state, neighbors)
        if bc_fn is not None:
            state = bc_fn(state)
        return state, neighbors

    advance._sph_fusable = _fusable
    # what an ensemble needs to issue this step itself (see ensemble.py)
    advance._sph_solver = solver
    advance._sph_tvf = tvf
    return advance

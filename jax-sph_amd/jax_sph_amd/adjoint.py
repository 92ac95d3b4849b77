"""Differentiable rollouts of the SPH step with a hand-written adjoint.

The reference gets gradients of a rollout with ``jax.grad`` through
``si_euler`` (``integrator.py:22-56``) and ``WCSPH.forward``
(``solver.py:705-949``).  Here the transpose of one step is a library call,
``sph_step_vjp``: two pair sweeps over the neighbour list of the step's new
positions.  :func:`rollout` wraps an ``n_steps`` rollout in a
``torch.autograd.Function``:

* the forward pass runs the fused step and keeps the state every
  ``checkpoint_every`` steps;
* the backward pass recomputes each segment from its checkpoint (the same
  kernels on exact pair lists, so bit for bit the forward trajectory), then
  applies the step transpose in reverse.

Memory is O(n_steps / k + k) states instead of O(n_steps).

Differentiated configuration (other settings raise ``NotImplementedError``):
standard SPH (``solver="SPH"``), density summation, no walls, no heat
conduction, no artificial viscosity, ``tvf = 0``, Quintic or Wendland-C2
kernel, and a constant body force.  Gradients flow to the initial velocity
and positions.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .integrator import si_euler
from .jax_md.partition import NeighborList
from .solver import WCSPH

_f = np.float32


def _check(solver: WCSPH, tvf: float):
    why = []
    if solver.solver != "SPH":
        why.append("solver must be 'SPH'")
    if solver.is_rho_evol:
        why.append("density evolution must be summation")
    if solver.is_bc_trick:
        why.append("walls are not differentiated")
    if solver.is_heat_conduction:
        why.append("heat conduction is not differentiated")
    if solver.artificial_alpha != 0.0:
        why.append("artificial viscosity is not differentiated")
    if tvf != 0.0:
        why.append("tvf must be 0")
    if solver._kernel_fn.name not in ("QSK", "WC2K"):
        why.append("kernel must be QSK or WC2K")
    if solver.g_const is None:
        why.append("the body force must be constant")
    if why:
        raise NotImplementedError("differentiable rollout: " + "; ".join(why))


class _Runner:
    """Static part of a differentiable rollout: solver, exact neighbour search,
    the fused step, and the step transpose."""

    def __init__(self, solver: WCSPH, neighbor_fn, shift_fn, dt: float, n_steps: int,
                 checkpoint_every: int, base_state: Dict[str, torch.Tensor]):
        self.solver = solver
        self.neighbor_fn = neighbor_fn
        self.dt = float(_f(dt))
        self.n_steps = int(n_steps)
        self.k = max(1, int(checkpoint_every))
        self.base = solver.state_in(base_state)
        self.advance = si_euler(0.0, solver.forward_wrapper(), shift_fn, None, None)
        self.prm = solver.params(tvf_stress=False)
        self.nb: Optional[NeighborList] = None
        self.reallocations = 0

    def _list_at(self, r):
        if self.nb is None:
            self.nb = self.neighbor_fn.allocate(r)
        else:
            self.nb = self.nb.update(r)
        return self.nb

    def _steps(self, st, nb, count, keep):
        """`count` fused steps from st; returns the last state and list (and the
        states after each step when keep); re-allocates and redoes the segment
        if its list overflowed (the error bits are sticky)."""
        while True:
            out, s, b = [], st, nb
            for _ in range(count):
                s, b = self.advance(self.dt, s, b)
                if keep:
                    out.append(s)
            if not bool(b.did_buffer_overflow):
                return s, b, out
            self.reallocations += 1
            self.nb = nb = self.neighbor_fn.allocate(st["r"])

    def forward(self, u0, r0):
        st = dict(self.base)
        st["u"], st["v"], st["r"] = u0, u0, r0
        nb = self._list_at(r0)
        ckpts = [st]
        done = 0
        while done < self.n_steps:
            c = min(self.k, self.n_steps - done)
            st, nb, _ = self._steps(st, nb, c, keep=False)
            done += c
            if done < self.n_steps:
                ckpts.append(st)
        self.nb = nb
        return st, ckpts

    def backward(self, ckpts, rb, ub):
        ab = torch.zeros_like(ub)
        bufs = [torch.empty_like(ub) for _ in range(3)]
        L = _lib.lib()
        for seg in reversed(range(len(ckpts))):
            start = seg * self.k
            count = min(self.k, self.n_steps - start)
            st0 = ckpts[seg]
            nb = self._list_at(st0["r"])
            _, _, states = self._steps(st0, nb, count, keep=True)
            for n in reversed(range(count)):
                nxt = states[n]
                # the exact pairs of the step's new positions
                nbn = self._list_at(nxt["r"])
                ws = nbn.workspace
                cin = self.solver.c_state_in(nxt)
                bar_in = _lib.CotangentIn(rb.data_ptr(), ub.data_ptr(), ab.data_ptr())
                rb2, ub2, ab2 = bufs
                bar_out = _lib.CotangentOut(rb2.data_ptr(), ub2.data_ptr(), ab2.data_ptr())
                _lib.check(L.sph_step_vjp(ws.handle, nbn.bank, ctypes.byref(self.prm), self.dt,
                                          ctypes.byref(cin), ctypes.byref(bar_in),
                                          ctypes.byref(bar_out), _lib.stream()),
                           "sph_step_vjp")
                bufs = [rb, ub, ab]
                rb, ub, ab = rb2, ub2, ab2
        return rb, ub


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, r0, runner: _Runner):
        st, ckpts = runner.forward(_lib.as_f32(u0.detach()), _lib.as_f32(r0.detach()))
        ctx.runner = runner
        ctx.ckpts = ckpts
        ctx.final = st
        return st["u"].clone(), st["r"].clone()

    @staticmethod
    def backward(ctx, gu, gr):
        runner = ctx.runner
        ub = _lib.as_f32(gu if gu is not None else torch.zeros_like(ctx.final["u"])).clone()
        rb = _lib.as_f32(gr if gr is not None else torch.zeros_like(ctx.final["r"])).clone()
        rb0, ub0 = runner.backward(ctx.ckpts, rb, ub)
        return ub0, rb0, None


def rollout(state: Dict[str, torch.Tensor], solver: WCSPH, neighbor_fn, shift_fn, dt: float,
            n_steps: int, tvf: float = 0.0, checkpoint_every: int = 10
            ) -> Tuple[torch.Tensor, torch.Tensor, "_Runner"]:
    """``n_steps`` si_euler steps from ``state``; returns the final (u, r).

    ``state["u"]`` and ``state["r"]`` may require grad: the returned tensors
    are differentiable with respect to them through the hand-written adjoint.
    ``neighbor_fn`` must build exact lists (``skin=0``), so the backward pass
    can recompute the forward trajectory exactly."""
    _check(solver, tvf)
    if getattr(getattr(neighbor_fn, "search", None), "skin", 0.0):
        raise ValueError("the differentiable rollout needs a neighbor_list with skin=0")
    runner = _Runner(solver, neighbor_fn, shift_fn, dt, n_steps, checkpoint_every, state)
    u, r = _Rollout.apply(state["u"], state["r"], runner)
    return u, r, runner

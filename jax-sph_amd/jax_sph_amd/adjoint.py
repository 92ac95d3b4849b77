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
standard SPH (``solver="SPH"``) as below, or the Riemann solver
(``solver="RIE"``, solver.py:108-164, 316-401) with density summation or its
continuity, any ``tvf``, with or without its wall condition (no-slip or
free-slip), its walls held by the case hook (the dam break as the survey's
benchmark runs it), or delta-SPH (``solver="DELTA"``) with density summation:
standard SPH plus its velocity diffusion (solver.py:259-311);
Quintic or Wendland-C2 kernel, a constant body force or (through :func:`rollout`) a
position-dependent one, with or without the
transport-velocity formulation (any ``tvf``: the stress ``A = rho u (v - u)``
and the background-pressure acceleration, solver.py:199-256), with or without
artificial viscosity (solver.py:404-428); density summation or density
continuity (``density_evolution``, solver.py:24-30; the density is then
carried from step to step, and checkpointed; with or without the density
renormalisation of solver.py:167-173), each with
or without the generalized wall boundary condition for no-slip or free-slip
walls (``is_bc_trick``, solver.py:431-530; free-slip walls at rest, whose
normals are those of the state).  The dam break as its case file runs
it (continuity, wall condition, artificial viscosity, gravity) is covered.
Heat conduction may be on: the temperature enters neither the EOS nor the
forces (solver.py:832-850), so the outputs u and r do not depend on it.  Wall
particles may be held by the case's boundary hook (``bc_fn``, a
``case_setup.CaseBC``: the lid-driven cavity as the reference's case file
runs it, and as its gradient check runs it, notebooks/iclr24_grads.ipynb).
Gradients flow to the initial velocity and positions.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .integrator import _bc_spec, si_euler
from .jax_md.partition import NeighborList
from .solver import WCSPH
from .utils import Tag, wall_tags

_f = np.float32


def _check(solver: WCSPH, tvf: float, bc_fn=None, graph: bool = False, state=None):
    why = []
    if solver.is_bc_trick and solver.is_free_slip and state is not None and \
            bool((torch.as_tensor(state["tag"]) == int(Tag.MOVING_WALL)).any()):
        why.append("free-slip moving walls (their normals follow them, case_setup.py:208-218) "
                   "are not differentiated")
    if bc_fn is not None and _bc_spec(bc_fn) is None:
        why.append("bc_fn must be a case boundary hook (case_setup.CaseBC)")
    if solver.solver not in ("SPH", "RIE", "DELTA"):
        why.append("solver must be 'SPH', 'RIE' or 'DELTA'")
    if solver.solver == "DELTA" and solver.is_rho_evol:
        why.append("delta-SPH is differentiated with density summation only (not the "
                   "density diffusion of its continuity, solver.py:33-105)")
    if solver.solver == "RIE" and state is not None:
        # the wall rows of the Riemann acceleration are not transposed: every
        # wall must be held by the case hook (its accelerations are constants)
        tag = torch.as_tensor(state["tag"])
        walls = torch.isin(tag, torch.as_tensor(wall_tags, device=tag.device))
        kind = getattr(bc_fn, "kind", 0) if bc_fn is not None else 0
        held = {1: (int(Tag.SOLID_WALL), int(Tag.MOVING_WALL)), 2: (int(Tag.SOLID_WALL),),
                3: tuple(int(t) for t in wall_tags)}.get(kind, ())
        if bool(walls.any()) and not bool(
                torch.isin(tag[walls], torch.as_tensor(held, dtype=tag.dtype,
                                                       device=tag.device)).all()):
            why.append("the Riemann solver's walls must be held by the case hook")
    if solver._kernel_fn.name not in ("QSK", "WC2K"):
        why.append("kernel must be QSK or WC2K")
    if solver.g_const is None and graph:
        why.append("the graph-replayed rollout needs a constant body force (rollout() "
                   "differentiates a position-dependent one)")
    if why:
        raise NotImplementedError("differentiable rollout: " + "; ".join(why))


def _tvf_coef(tvf: float, dt: float) -> np.float32:
    """c of v' = u' + c dvdt, rounded as the fused step rounds it."""
    return _f(_f(tvf * 0.5) * _f(dt))


def _ptr(t):
    return t.data_ptr() if t is not None else None


def _cot_in(c: Dict[str, torch.Tensor]):
    """sph_cotangent_in of the carried cotangents c (keys of _carried); the
    cotangent of v' is zero (the next step recomputes v from u and dvdt)."""
    return _lib.CotangentIn(_ptr(c["r"]), _ptr(c["u"]), _ptr(c["dudt"]), None,
                            _ptr(c.get("dvdt")), _ptr(c.get("rho")))


def _cot_out(c: Dict[str, torch.Tensor]):
    return _lib.CotangentOut(_ptr(c["r"]), _ptr(c["u"]), _ptr(c["dudt"]),
                             _ptr(c.get("dvdt")), _ptr(c.get("rho")))


class _Runner:
    """Static part of a differentiable rollout: solver, exact neighbour search,
    the fused step, and the step transpose."""

    def __init__(self, solver: WCSPH, neighbor_fn, shift_fn, dt: float, n_steps: int,
                 checkpoint_every: int, base_state: Dict[str, torch.Tensor], tvf: float = 0.0,
                 bc_fn=None):
        self.solver = solver
        self.neighbor_fn = neighbor_fn
        self.dt = float(_f(dt))
        self.n_steps = int(n_steps)
        self.k = max(1, int(checkpoint_every))
        self.base = solver.state_in(base_state)
        self.tvf = float(tvf)
        self.carried = _carried(self.tvf, solver.is_rho_evol)
        self.advance = si_euler(self.tvf, solver.forward_wrapper(), shift_fn, bc_fn, None)
        self.prm = solver.params(tvf_stress=self.tvf != 0.0)
        self.ctvf = float(_tvf_coef(self.tvf, self.dt))
        self.bc = _bc_spec(bc_fn)
        self.bc_fn = bc_fn
        self.nb: Optional[NeighborList] = None
        self.reallocations = 0
        self._kept: Optional[torch.Tensor] = None

    def _body_force_pullback(self, st, cur):
        """A position-dependent body force adds g(r') to the step's dudt'
        (solver.py:936, before the case hook), so r' also receives
        (dg/dr')^T of the cotangent of dudt' where the hook keeps dudt'.  The
        derivative is taken by autograd through the caller's g_ext_fn (a torch
        function of r); a force built from comparisons (torch.where, as the
        Poiseuille cases' are) has none, as under jax.grad."""
        r = st["r"].detach().clone().requires_grad_(True)
        with torch.enable_grad():
            gv = self.solver.g_ext_fn(r)
        if not (isinstance(gv, torch.Tensor) and gv.requires_grad):
            return
        a = cur["dudt"]
        if self.bc_fn is not None:
            if self._kept is None:   # 1 where the hook leaves dudt alone (tags are fixed)
                probe = dict(st)
                probe["dudt"] = torch.ones_like(st["dudt"])
                probe["dvdt"] = torch.ones_like(st["dvdt"])
                self._kept = self.bc_fn(probe)["dudt"].detach().clone()
            a = a * self._kept
        (jr,) = torch.autograd.grad(gv, r, grad_outputs=a.to(gv.dtype), allow_unused=True)
        if jr is not None:
            cur["r"] += jr.to(cur["r"].dtype)

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
        # cotangents of the carried state; the rollout's outputs are u and r
        cur = {key: torch.zeros_like(ckpts[0][key]) for key in self.carried}
        cur["r"], cur["u"] = rb, ub
        spare = {key: torch.empty_like(v) for key, v in cur.items()}
        L = _lib.lib()
        for seg in reversed(range(len(ckpts))):
            start = seg * self.k
            count = min(self.k, self.n_steps - start)
            st0 = ckpts[seg]
            nb = self._list_at(st0["r"])
            _, _, states = self._steps(st0, nb, count, keep=True)
            for n in reversed(range(count)):
                if self.solver.is_rho_evol:
                    # the transpose needs the step's records: the continuity
                    # update starts from the density of the state before, which
                    # state n+1 does not hold, so the step runs again from there
                    prev = states[n - 1] if n > 0 else st0
                    nxt, nbn = self.advance(self.dt, prev, self._list_at(prev["r"]))
                    ready = 1
                else:
                    nxt = states[n]
                    # the exact pairs of the step's new positions
                    nbn = self._list_at(nxt["r"])
                    ready = 0
                ws = nbn.workspace
                g = self.solver.g_ext_tensor(nxt["r"])   # None for a constant force
                if g is not None:
                    self._body_force_pullback(nxt, cur)
                cin = self.solver.c_state_in(nxt, g)
                bar_in = _cot_in(cur)
                bar_out = _cot_out(spare)
                _lib.check(L.sph_step_vjp(ws.handle, nbn.bank, ctypes.byref(self.prm), self.dt,
                                          self.ctvf, ctypes.byref(self.bc), ctypes.byref(cin),
                                          ctypes.byref(bar_in), ctypes.byref(bar_out), ready,
                                          _lib.stream()), "sph_step_vjp")
                cur, spare = spare, cur
        return cur["r"], cur["u"]


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


EVOLVING = ("r", "u", "v", "dudt", "dvdt", "rho", "p", "drhodt")


# What one step of the differentiated configuration reads from the state it
# starts from: positions, velocity and the accelerations (v = u + c dvdt is
# recomputed; with density summation so are the fluid density and pressure,
# by the summation and the EOS).  Without tvf, c = 0 and dvdt is not read;
# with density continuity the step advances the density it is given.
def _carried(tvf: float, continuity: bool = False) -> Tuple[str, ...]:
    keys = ("r", "u", "dudt", "dvdt") if tvf != 0.0 else ("r", "u", "dudt")
    return keys + ("rho",) if continuity else keys


class _Flat:
    """The carried fields of a state in one flat buffer, so that a checkpoint
    is one copy; `views` maps each field to its (N, d) or (N,) part."""

    def __init__(self, fields: Dict[str, torch.Tensor], keys: Tuple[str, ...]):
        self.keys = keys
        self.shapes = [tuple(fields[k].shape) for k in keys]
        self.buf = torch.cat([fields[k].reshape(-1) for k in keys]).contiguous()

    def views(self, buf: torch.Tensor) -> Dict[str, torch.Tensor]:
        out, off = {}, 0
        for key, shape in zip(self.keys, self.shapes):
            n = int(np.prod(shape))
            out[key] = buf[off:off + n].view(shape)
            off += n
        return out


class GraphRollout:
    """The same differentiable rollout with both passes as HIP graphs.

    Each of the ``k = checkpoint_every`` steps of a segment has a workspace of
    its own.  The exact pair list that step builds for its new positions
    therefore survives the segment, and the reverse sweep needs no rebuild.
    ``workspace_sets`` such sets rotate over the segments.  Graphs:

    * A: one forward segment of ``k`` fused steps, in place;
    * R[s]: recompute a segment from its checkpoint into set s, keeping each
      step's state;
    * V[s]: the ``k`` step transposes over set s, updating the cotangent
      buffers in place.

    A recompute needs only its checkpoint.  The backward pass therefore runs
    the recomputes of later-reversed segments on ``side_streams`` streams
    while the calling stream applies V, segment after segment.  A set is
    reused only after its V has run.  ``n_steps`` must be a multiple of
    ``k``.  The gradients equal :func:`rollout`'s bit for bit.

    With as many sets as segments, the forward pass itself runs R[s] for
    segment s, so every step's state and pair list stays in place, and the
    backward pass of that forward only replays the V graphs: memory for the
    recompute.  ``workspace_sets=None`` takes that route when all sets fit in
    half of the free device memory, else 4 sets up to 100k particles and 2
    above.  The workspaces hold ``list_margin`` times the largest pair list
    of the start per particle; a list that outgrows that raises the overflow
    error below (a larger margin, or :func:`rollout`, which re-allocates).
    """

    def __init__(self, state: Dict[str, torch.Tensor], solver: WCSPH, neighbor_fn, shift_fn,
                 dt: float, n_steps: int, tvf: float = 0.0, checkpoint_every: int = 10,
                 workspace_sets: Optional[int] = None, side_streams: int = 3, bc_fn=None,
                 list_margin: float = 1.5):
        _check(solver, tvf, bc_fn, graph=True, state=state)
        if getattr(getattr(neighbor_fn, "search", None), "skin", 0.0):
            raise ValueError("the differentiable rollout needs a neighbor_list with skin=0")
        self.k = int(checkpoint_every)
        if n_steps % self.k:
            raise ValueError("n_steps must be a multiple of checkpoint_every")
        self.n_steps = int(n_steps)
        self.solver = solver
        self.dt = float(_f(dt))
        self.tvf = float(tvf)
        self.ctvf = float(_tvf_coef(self.tvf, self.dt))
        self.prm = solver.params(tvf_stress=self.tvf != 0.0)
        self.advance = si_euler(self.tvf, solver.forward_wrapper(), shift_fn, bc_fn, None)
        self.bc = _bc_spec(bc_fn)
        carried = _carried(self.tvf, solver.is_rho_evol)
        self.carried = carried
        base = solver.state_in(state)
        n_seg = self.n_steps // self.k
        # Each step keeps its list in a workspace of its own, so the pair
        # buffers are sized from the largest list at the start (times
        # list_margin) rather than for 3^d x cell capacity candidates per
        # particle; a list that outgrows them raises the overflow bit.
        probe = neighbor_fn.allocate(base["r"])
        width = int(np.ceil(int(probe.info[3]) * float(list_margin))) + 16
        del probe

        def alloc():
            return neighbor_fn.allocate(base["r"], ell_width=width)
        self.sets = [[alloc() for _ in range(self.k)]]
        if workspace_sets is None:
            # one set: its workspaces, plus the k states its recompute keeps
            # and the graphs' scratch, both counted as much again
            per_set = sum(nb.workspace.bytes for nb in self.sets[0])
            per_set += 2 * self.k * sum(v.numel() * v.element_size() for v in base.values())
            free, _ = torch.cuda.mem_get_info()
            if n_seg * per_set <= free // 2:
                workspace_sets = n_seg
            else:   # concurrency pays at small sizes; memory rules large ones
                workspace_sets = 4 if base["r"].shape[0] <= 100_000 else 2
        self.n_sets = max(1, min(int(workspace_sets), n_seg))
        # every segment keeps its steps in a set of its own: no recompute
        self.keep_all = self.n_sets == n_seg
        self.sets += [[alloc() for _ in range(self.k)] for _ in range(self.n_sets - 1)]
        # carried fields live in one flat buffer: a checkpoint is one copy
        self.flat = _Flat(base, carried)
        self.base_buf = self.flat.buf
        self.S_buf = self.base_buf.clone()
        self.S = {key: v.clone() for key, v in base.items()}   # forward segment state
        self.S.update(self.flat.views(self.S_buf))
        self.C_bufs, self.C = [], []
        for _ in range(self.n_sets):
            cb = self.base_buf.clone()
            c = {key: v.clone() for key, v in base.items()}
            c.update(self.flat.views(cb))
            self.C_bufs.append(cb)
            self.C.append(c)
        # cotangents of the carried fields
        self.bar = {key: torch.zeros_like(base[key]) for key in carried}
        self.bits = torch.zeros((), dtype=torch.int64, device=base["u"].device)
        self.sides = [torch.cuda.Stream() for _ in range(max(1, int(side_streams)))]
        self.generation = 0
        torch.cuda.synchronize()
        self.gA = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.gA):
            st = self.S
            bits = None
            for j in range(self.k):
                st, nb = self.advance(self.dt, st, self.sets[0][j])
                b = nb.info[0]
                bits = b if bits is None else torch.bitwise_or(bits, b)
            for key in EVOLVING:
                self.S[key].copy_(st[key])
            self.bits.copy_(torch.bitwise_or(self.bits, bits))
        L = _lib.lib()
        self.gR, self.gV, self._steps = [], [], []
        for s in range(self.n_sets):
            gR = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gR):
                st, steps = self.C[s], []
                bits = None
                for j in range(self.k):
                    st, nb = self.advance(self.dt, st, self.sets[s][j])
                    steps.append((st, nb))
                    b = nb.info[0]
                    bits = b if bits is None else torch.bitwise_or(bits, b)
                self.bits.copy_(torch.bitwise_or(self.bits, bits))
            gV = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gV):
                cur = dict(self.bar)
                for st_j, nb_j in reversed(steps):
                    out = {key: torch.empty_like(v) for key, v in cur.items()}
                    bar_in = _cot_in(cur)
                    bar_out = _cot_out(out)
                    _lib.check(L.sph_step_vjp(nb_j.workspace.handle, nb_j.bank,
                                              ctypes.byref(self.prm), self.dt, self.ctvf,
                                              ctypes.byref(self.bc),
                                              ctypes.byref(self.solver.c_state_in(st_j)),
                                              ctypes.byref(bar_in), ctypes.byref(bar_out),
                                              1,  # records left by this step's own R pass
                                              _lib.stream()), "sph_step_vjp")
                    cur = out
                for key, dst in self.bar.items():
                    dst.copy_(cur[key])
            self.gR.append(gR)
            self.gV.append(gV)
            self._steps.append(steps)   # keeps the recomputed states alive

    def forward(self, u0, r0):
        """Runs the rollout; returns (u_final, r_final, checkpoints)."""
        if self.keep_all:
            return self._forward_kept(u0, r0)
        self.S_buf.copy_(self.base_buf)
        self.S["u"].copy_(u0)
        self.S["v"].copy_(u0)
        self.S["r"].copy_(r0)
        self.bits.zero_()
        ckpts = []
        n = self.n_steps // self.k
        for _ in range(n - 1):
            ckpts.append(self.S_buf.clone())
            self.gA.replay()
        # the last segment runs as a recompute graph, so its states and pair
        # lists are already in place when the backward pass starts
        ckpts.append(self.S_buf.clone())
        s = (n - 1) % self.n_sets
        self.C_bufs[s].copy_(self.S_buf)
        self.gR[s].replay()
        if int(self.bits) & 3:
            raise RuntimeError("neighbour-list overflow inside the graph rollout; use "
                               "adjoint.rollout (re-allocates) or a larger capacity")
        self.generation += 1
        last = self._steps[s][-1][0]
        return last["u"].clone(), last["r"].clone(), ckpts, self.generation

    def _forward_kept(self, u0, r0):
        """forward() with a set per segment: segment s runs R[s] from the end
        state of segment s - 1, and all of them stay in place."""
        carried = self.carried
        n = self.n_steps // self.k
        self.bits.zero_()
        ckpts = []
        for seg in range(n):
            c = self.C[seg]
            if seg == 0:
                self.C_bufs[0].copy_(self.base_buf)
                c["u"].copy_(u0)
                c["v"].copy_(u0)
                c["r"].copy_(r0)
            else:
                last = self._steps[seg - 1][-1][0]
                for key in carried:
                    c[key].copy_(last[key])
                c["v"].copy_(last["v"])
            ckpts.append(self.C_bufs[seg].clone())   # for a backward after a later forward
            self.gR[seg].replay()
        if int(self.bits) & 3:
            raise RuntimeError("neighbour-list overflow inside the graph rollout; use "
                               "adjoint.rollout (re-allocates) or a larger capacity")
        self.generation += 1
        last = self._steps[n - 1][-1][0]
        return last["u"].clone(), last["r"].clone(), ckpts, self.generation

    def backward(self, ckpts, ub_final, rb_final, generation=None):
        """Cotangents of (r0, u0).  When no other forward ran since the one
        that made `ckpts` (same generation), its last segment (with a set per
        segment: every segment) is still in place and is not recomputed."""
        main = torch.cuda.current_stream()
        for key, b in self.bar.items():
            if key == "r":
                b.copy_(rb_final)
            elif key == "u":
                b.copy_(ub_final)
            else:
                b.zero_()
        n = len(ckpts)
        if self.keep_all and generation == self.generation:
            for seg in reversed(range(n)):
                self.gV[seg].replay()
            return self.bar["r"].clone(), self.bar["u"].clone()
        ev_r = [torch.cuda.Event() for _ in range(n)]
        ev_v = [torch.cuda.Event() for _ in range(n)]
        for side in self.sides:
            side.wait_stream(main)
        ns = self.n_sets

        def recompute(seg):
            s = seg % ns
            side = self.sides[seg % len(self.sides)]
            with torch.cuda.stream(side):
                if seg + ns < n:
                    side.wait_event(ev_v[seg + ns])   # set s was last read by V(seg + ns)
                self.C_bufs[s].copy_(ckpts[seg])
                self.gR[s].replay()
                ev_r[seg].record(side)
            ckpts[seg].record_stream(side)

        if generation == self.generation:
            ev_r[n - 1].record(main)    # the forward left segment n-1 in its set
        else:
            recompute(n - 1)
        for seg in range(n - 2, max(-1, n - 1 - ns), -1):
            recompute(seg)              # sets not yet used by any V
        for seg in reversed(range(n)):
            main.wait_event(ev_r[seg])
            self.gV[seg % ns].replay()
            ev_v[seg].record(main)
            if seg - ns >= 0:
                recompute(seg - ns)     # reuses the set V(seg) just released
        return self.bar["r"].clone(), self.bar["u"].clone()

    def __call__(self, u0: torch.Tensor, r0: torch.Tensor):
        """Differentiable (u_final, r_final) of the rollout from (u0, r0)."""
        return _GraphFn.apply(u0, r0, self)


class _GraphFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, r0, runner: GraphRollout):
        u, r, ckpts, gen = runner.forward(_lib.as_f32(u0.detach()), _lib.as_f32(r0.detach()))
        ctx.runner, ctx.ckpts, ctx.generation = runner, ckpts, gen
        return u, r

    @staticmethod
    def backward(ctx, gu, gr):
        runner = ctx.runner
        ub = gu if gu is not None else torch.zeros_like(runner.S["u"])
        rb = gr if gr is not None else torch.zeros_like(runner.S["r"])
        rb0, ub0 = runner.backward(ctx.ckpts, _lib.as_f32(ub), _lib.as_f32(rb), ctx.generation)
        return ub0, rb0, None


def rollout(state: Dict[str, torch.Tensor], solver: WCSPH, neighbor_fn, shift_fn, dt: float,
            n_steps: int, tvf: float = 0.0, checkpoint_every: int = 10, bc_fn=None
            ) -> Tuple[torch.Tensor, torch.Tensor, "_Runner"]:
    """``n_steps`` si_euler steps from ``state``; returns the final (u, r).

    ``state["u"]`` and ``state["r"]`` may require grad: the returned tensors
    are differentiable with respect to them through the hand-written adjoint.
    ``neighbor_fn`` must build exact lists (``skin=0``), so the backward pass
    can recompute the forward trajectory exactly.  ``bc_fn`` is the case's
    boundary hook (a ``case_setup.CaseBC``), applied after every step."""
    _check(solver, tvf, bc_fn, state=state)
    if getattr(getattr(neighbor_fn, "search", None), "skin", 0.0):
        raise ValueError("the differentiable rollout needs a neighbor_list with skin=0")
    runner = _Runner(solver, neighbor_fn, shift_fn, dt, n_steps, checkpoint_every, state, tvf,
                     bc_fn)
    u, r = _Rollout.apply(state["u"], state["r"], runner)
    return u, r, runner

"""Ensembles of independent simulations, one launch per phase of the step.

The reference advances one simulation per ``simulate`` call
(``simulate.py:19-136``); an ensemble of runs there is a host loop over
independent simulations.  Small simulations leave most of an MI355X idle:
a 2,500-particle step is a dozen launches of a few workgroups each.
:class:`Ensemble` instead issues every phase of the fused ``si_euler`` step
(kick/drift, Verlet-skin rebuild test, binning, pair build, density, walls,
forces, counters) as ONE launch over all members (``sph_ensemble_step``).
Each member keeps its own workspace, skin lists, parameters and error block,
and its results are bitwise those of its own fused step.

Members alternate between two state buffer sets (A -> B, then B -> A), so the
step allocates nothing and :meth:`capture` records a fixed number of steps
into a HIP graph that replays them with one launch.  :meth:`autotune` times
the block layouts of the batched launches (member order, XCD-contiguous
workgroups) on graph-replayed steps and keeps the fastest; it restores the
state afterwards and makes every member rebuild its lists, so a run tuned
right after construction follows the untuned trajectory bit for bit.
:meth:`tuned_from_configs` also searches each group's Verlet-skin width (the
skin cells' size, i.e. the cell tile of the pair sweeps); members keep
equalling their own fused steps with the skin chosen for them.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from . import partition as _partition
from .case_setup import Cfg, load_case, make_solver
from .integrator import si_euler
from .jax_md.partition import NeighborList, Sparse
from .solver import STATE_KEYS

_f = np.float32

# state the fused step writes (the others are constant through a step)
_WRITTEN = ("r", "u", "v", "dudt", "dvdt", "rho", "p", "drhodt")


def group_key(cfg: Cfg) -> Tuple:
    """The solver settings that select a member's sweep variants (density
    mode, walls, force solver, transport velocity, kernel, extras): members
    with equal keys share the batched sweep launches of one group."""
    s = cfg.solver
    return (s.name, bool(s.density_evolution), bool(s.is_bc_trick), float(s.tvf) != 0.0,
            cfg.kernel.name, float(s.artificial_alpha) != 0.0, bool(s.heat_conduction))


def setup_member(cfg: Cfg, skin="auto", capacity_multiplier: float = 1.25,
                 skin_factor: Optional[float] = None):
    """One simulation set up as ``simulate`` does it (simulate.py:35-99): case
    state, solver, ``jaxmd_vmap`` neighbour list, fused ``si_euler``.
    ``skin_factor`` (optional) sets the Verlet skin to that multiple of the
    member's cutoff.  Returns ``(advance, state, neighbors, dt)``."""
    Case = load_case(cfg.case.source or cfg.config)
    (cfg, box, state, g_ext_fn, bc_fn, nw_fn, eos, key, disp, shift) = Case(cfg).initialize()
    solver = make_solver(cfg, disp, eos, g_ext_fn)
    if skin_factor is not None:
        skin = float(skin_factor) * float(solver._kernel_fn.cutoff)
    fn = _partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                  backend="jaxmd_vmap", capacity_multiplier=capacity_multiplier,
                                  mask_self=False, format=Sparse, skin=skin)
    nb = fn.allocate(state["r"])
    adv = si_euler(cfg.solver.tvf, solver.forward_wrapper(), shift, bc_fn, nw_fn)
    return adv, state, nb, cfg.solver.dt


class _Member:
    def __init__(self, k, advance, state, neighbors: NeighborList, dt: float):
        fusable = getattr(advance, "_sph_fusable", None)
        spec = fusable(neighbors) if fusable is not None else None
        if spec is None:
            raise ValueError(f"member {k}: not a fused step (si_euler of WCSPH.forward_wrapper, "
                             "periodic shift, case BC, constant body force, dr_threshold 0)")
        search = neighbors._search
        if search.step_bank(neighbors) != _lib.SKIN_BANK:
            raise ValueError(f"member {k}: the neighbour list has no Verlet-skin lists "
                             "(jaxmd_vmap backend with skin enabled)")
        neighbors._check_live()
        solver, tvf = advance._sph_solver, advance._sph_tvf
        self.solver = solver
        self.ws = neighbors.workspace
        self.n = self.ws.n
        s = solver.state_in(state)
        self.buf = [{key: v.clone() for key, v in s.items()},
                    {key: v.clone() for key, v in s.items()}]
        self.info = [neighbors.info.clone(), torch.zeros_like(neighbors.info)]
        self.max_occupancy = neighbors.max_occupancy
        self.dt = _f(dt)
        self.ctvf = _f(_f(tvf * 0.5) * self.dt)
        self.prm = solver.params(tvf_stress=(tvf != 0.0))
        self.spec = spec
        # the lists now belong to the ensemble: the caller's NeighborList is stale
        self.ws.generation[_lib.SKIN_BANK] += 1
        # ctypes structs of the two directions (kept alive with the member)
        self.c = []
        for src in (0, 1):
            a, b = self.buf[src], self.buf[1 - src]
            outs = {key: b[key] for key in ("u", "v", "dudt", "dvdt", "rho", "p", "drhodt")}
            if solver.is_heat_conduction:
                outs["T"], outs["dTdt"] = b["T"], b["dTdt"]
            cin, cout = solver.c_state_in(a), solver.c_state_out(outs)
            self.c.append((cin, cout))

    def member_struct(self, src: int) -> _lib.EnsembleMember:
        cin, cout = self.c[src]
        m = _lib.EnsembleMember()
        m.ws = self.ws.handle
        m.prm = ctypes.pointer(self.prm)
        m.dt = float(self.dt)
        m.tvf_coef = float(self.ctvf)
        m.bc = ctypes.pointer(self.spec)
        m.in_ = ctypes.pointer(cin)
        m.r_out = self.buf[1 - src]["r"].data_ptr()
        m.out = ctypes.pointer(cout)
        m.prev_info = self.info[src].data_ptr()
        m.max_occupancy = int(self.max_occupancy)
        m.info = self.info[1 - src].data_ptr()
        return m


class Ensemble:
    """Independent fused-step simulations advanced together.

    ``members``: ``(advance, state, neighbors, dt)`` per simulation, with
    ``advance`` from :func:`integrator.si_euler` over ``WCSPH.forward_wrapper()``
    and ``neighbors`` an allocated ``jaxmd_vmap`` list with Verlet-skin lists
    (the default ``skin="auto"``).  All members have the same dimension.
    """

    def __init__(self, members: Iterable[Tuple], order: int = 1, xcd: bool = True):
        self.members: List[_Member] = [_Member(k, *m) for k, m in enumerate(members)]
        if not self.members:
            raise ValueError("an ensemble needs at least one member")
        if len({m.solver.dim for m in self.members}) != 1:
            raise ValueError("every member must have the same dimension")
        self.parity = 0      # members' current state is buf[parity]
        self.steps = 0
        self._graph = None
        self._graph_steps = 0
        self.handles = []
        L = _lib.lib()
        for src in (0, 1):
            arr = (_lib.EnsembleMember * len(self.members))(
                *[m.member_struct(src) for m in self.members])
            h = ctypes.c_void_p()
            _lib.check(L.sph_ensemble_create(arr, len(self.members), ctypes.byref(h)),
                       "sph_ensemble_create")
            self.handles.append(h)
        self.layout = (None, None)
        self.set_layout(order, xcd)

    # ---------------------------------------------------------------- build
    @classmethod
    def from_configs(cls, cfgs: Sequence[Cfg], skin="auto", capacity_multiplier: float = 1.25,
                     order: int = 1, xcd: bool = True,
                     skin_factors: Optional[Sequence[float]] = None) -> "Ensemble":
        """Sets up each config as ``simulate`` does (case, solver, neighbour
        list, ``si_euler``) and builds the ensemble of them.  ``skin_factors``
        (optional, one per config) sets each member's Verlet skin to that
        multiple of its cutoff."""
        factors = list(skin_factors) if skin_factors is not None else [None] * len(cfgs)
        members = [setup_member(cfg, skin, capacity_multiplier, f)
                   for cfg, f in zip(cfgs, factors)]
        ens = cls(members, order=order, xcd=xcd)
        ens.skin_factors = factors
        return ens

    @classmethod
    def tuned_from_configs(cls, cfgs: Sequence[Cfg], factors: Sequence[float] = (0.05, 0.1, 0.15),
                           steps: int = 10, reps: int = 2,
                           capacity_multiplier: float = 1.25) -> Tuple["Ensemble", Dict]:
        """The ensemble of ``cfgs`` with its launch shapes autotuned: the
        Verlet skin of each group's members (the width of its skin cells, the
        "cell tile" of the pair sweeps: wider lists against rarer rebuilds),
        then the block layout (:meth:`autotune`).  The skins are searched
        jointly over ``factors`` (multiples of each member's cutoff), then
        group by group with the other groups at their best.  Each candidate
        is a freshly set-up ensemble timed on graph-replayed steps from the
        initial state; the one returned starts from the initial state too."""
        cfgs = list(cfgs)
        keys = [group_key(c) for c in cfgs]
        groups = sorted(set(keys), key=str)
        trials = {}

        def time_of(per_group):
            tag = tuple(per_group[g] for g in groups)
            if tag not in trials:
                ens = cls.from_configs(cfgs, capacity_multiplier=capacity_multiplier,
                                       skin_factors=[per_group[k] for k in keys])
                res = ens.autotune(steps=steps, reps=reps)
                trials[tag] = (min(res["ms_per_step"].values()), res["best"])
                del ens
            return trials[tag][0]

        best = {g: factors[0] for g in groups}
        best_t = None
        for f in factors:
            cand = {g: f for g in groups}
            t = time_of(cand)
            if best_t is None or t < best_t:
                best, best_t = cand, t
        if len(groups) > 1:
            for g in groups:
                for f in factors:
                    cand = dict(best)
                    cand[g] = f
                    t = time_of(cand)
                    if t < best_t:
                        best, best_t = cand, t
        layout = trials[tuple(best[g] for g in groups)][1]
        ens = cls.from_configs(cfgs, capacity_multiplier=capacity_multiplier,
                               order=layout["order"], xcd=layout["xcd"],
                               skin_factors=[best[k] for k in keys])
        report = {"groups": [str(g) for g in groups],
                  "skin_factor_per_group": [best[g] for g in groups],
                  "layout": layout, "ms_per_step": best_t,
                  "trials_ms_per_step": {"/".join(f"{x:g}" for x in tag): t
                                         for tag, (t, _) in trials.items()}}
        return ens, report

    def __del__(self):
        try:
            L = _lib.lib()
            for h in getattr(self, "handles", []):
                L.sph_ensemble_destroy(h)
        except Exception:
            pass

    # ---------------------------------------------------------------- layout
    def set_layout(self, order: int, xcd: bool):
        """Block layout of the batched launches: member ``order`` 0 (given),
        1 (largest first) or 2 (blocks interleaved); ``xcd`` contiguous runs
        per XCD.  Invalidates a captured graph."""
        if (order, bool(xcd)) == self.layout:
            return
        L = _lib.lib()
        for h in self.handles:
            _lib.check(L.sph_ensemble_tune(h, b"order", int(order)), "sph_ensemble_tune")
            _lib.check(L.sph_ensemble_tune(h, b"xcd", int(bool(xcd))), "sph_ensemble_tune")
        self.layout = (order, bool(xcd))
        self._graph = None

    def info(self) -> Dict[str, int]:
        m, g, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib().sph_ensemble_info(self.handles[0], ctypes.byref(m),
                                                ctypes.byref(g), ctypes.byref(b)),
                   "sph_ensemble_info")
        return {"members": m.value, "groups": g.value, "lane_blocks": b.value,
                "particles": sum(x.n for x in self.members)}

    # ---------------------------------------------------------------- steps
    def _launch(self):
        _lib.check(_lib.lib().sph_ensemble_step(self.handles[self.parity], _lib.stream()),
                   "sph_ensemble_step")
        self.parity ^= 1

    def step(self, n: int = 1):
        """Advances every member by ``n`` fused steps (eager launches)."""
        for _ in range(n):
            self._launch()
        self.steps += n

    def capture(self, steps: int = 2):
        """Records ``steps`` (even) steps of every member into a HIP graph."""
        if steps < 2 or steps % 2:
            raise ValueError("capture an even number of steps (the buffers alternate)")
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        parity = self.parity
        with torch.cuda.graph(g):
            for _ in range(steps):
                self._launch()
        self.parity = parity  # capture launches nothing
        self._graph, self._graph_steps = g, steps

    def replay(self, n: int = 1):
        """Advances every member by ``n`` times the captured step count."""
        if self._graph is None:
            raise RuntimeError("no captured graph (capture() first; set_layout drops it)")
        for _ in range(n):
            self._graph.replay()
        self.steps += n * self._graph_steps

    # ---------------------------------------------------------------- state
    def state(self, k: int) -> Dict[str, torch.Tensor]:
        """Member k's current 16-key state (views of the ensemble's buffers)."""
        b = self.members[k].buf[self.parity]
        return {key: b[key] for key in STATE_KEYS}

    def error_info(self, k: int) -> torch.Tensor:
        """Member k's int64[8] error block after the last step."""
        return self.members[k].info[self.parity]

    def overflowed(self) -> List[int]:
        """Members whose lists overflowed (sticky error bits & 3)."""
        codes = torch.stack([m.info[self.parity][0] for m in self.members]).cpu()
        return [k for k, c in enumerate(codes.tolist()) if c & 3]

    @property
    def did_buffer_overflow(self) -> bool:
        return bool(self.overflowed())

    # ---------------------------------------------------------------- tuning
    def _snapshot(self):
        return (self.parity, self.steps,
                [({key: v.clone() for key, v in m.buf[self.parity].items()},
                  m.info[self.parity].clone()) for m in self.members])

    def _restore(self, snap):
        parity, steps, data = snap
        self.parity, self.steps = parity, steps
        for m, (buf, info) in zip(self.members, data):
            for key, v in buf.items():
                m.buf[parity][key].copy_(v)
            m.info[parity].copy_(info)
        torch.cuda.synchronize()
        for m in self.members:
            m.ws.tune("skin_rebuild", 1)

    def autotune(self, candidates: Optional[Sequence[Tuple[int, bool]]] = None,
                 steps: int = 10, reps: int = 3) -> Dict:
        """Times each block layout on ``reps`` replays of a ``steps``-step
        graph and keeps the fastest.  State and lists are restored after every
        trial (the members rebuild their skin lists on the next step)."""
        cands = list(candidates or [(0, False), (0, True), (1, False), (1, True), (2, True)])
        snap = self._snapshot()
        times = {}
        for order, xcd in cands:
            self.set_layout(order, xcd)
            self.capture(steps)
            self.replay(1)  # warm-up
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record()
            self.replay(reps)
            t1.record()
            t1.synchronize()
            times[(order, bool(xcd))] = t0.elapsed_time(t1) / (reps * steps)
            self._graph = None
            self._restore(snap)
        best = min(times, key=times.get)
        self.set_layout(*best)
        self._restore(snap)
        return {"best": {"order": best[0], "xcd": best[1]},
                "ms_per_step": {f"order={o},xcd={int(x)}": t for (o, x), t in times.items()}}

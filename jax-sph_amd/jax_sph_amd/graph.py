"""HIP-graph replay of fused si_euler steps.

At small particle counts a step is a dozen short kernels, and the host-side
launch cost dominates.  :class:`StepGraph` captures ``k`` fused steps once
(``torch.cuda.graph`` records every kernel the library launches on the
capture stream) and replays them as one graph launch.  Each replay advances
the held state by ``k`` steps in place: the graph ends by copying the new
state into its input buffers and the new error block into the one the first
step reads, so the error bits stay sticky from replay to replay.

The step itself allocates nothing and makes no host decisions, so the graph
replays exactly the eager computation.  The Verlet-skin rebuild test runs on
the device.  The library records no memset node: fills are kernels.  The
results are bitwise those of the eager steps (tested).  A list overflow
inside a replay shows in :attr:`did_buffer_overflow`; the caller then
re-allocates and redoes the steps eagerly, as after any overflow.
"""
from __future__ import annotations

from typing import Dict

import torch

from .jax_md.partition import NeighborList


class StepGraph:
    def __init__(self, advance, state: Dict[str, torch.Tensor], neighbors: NeighborList,
                 dt: float, steps_per_replay: int = 1):
        if getattr(advance, "_sph_fusable", None) is None or \
                advance._sph_fusable(neighbors) is None:
            raise ValueError("StepGraph needs the fused step (see integrator.si_euler)")
        self.k = int(steps_per_replay)
        if self.k < 1:
            raise ValueError("steps_per_replay must be >= 1")
        self.state = {key: v.clone() for key, v in state.items()}
        self._info0 = neighbors.info
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            st, nb = self.state, neighbors
            for _ in range(self.k):
                st, nb = advance(dt, st, nb)
            for key, v in self.state.items():
                if st[key] is not v:
                    v.copy_(st[key])
            self._info0.copy_(nb.info)
        self.neighbors = nb
        self.replays = 0

    def replay(self, n: int = 1):
        """Advances the held state by n * steps_per_replay steps."""
        for _ in range(n):
            self.graph.replay()
        self.replays += n

    @property
    def did_buffer_overflow(self) -> bool:
        return bool(self._info0[0] & 3)

"""Weakly compressible SPH right-hand side (drop-in for ``jax_sph/solver.py:613-951``).

``WCSPH(...).forward_wrapper()`` returns ``forward(state, neighbors)`` with the
reference's 16-key state dict.  The whole RHS runs in the HIP library
(``sph_forward``) as pair sweeps over the neighbour list:

* a density sweep: summation, SPH / Riemann / delta-SPH continuity, plus the
  Riemann wall velocity.  Delta-SPH first needs two sweeps for its
  renormalisation matrices and density gradients;
* optional density renormalisation;
* the wall sweep: generalized wall BC, or the Riemann heat BC;
* one force sweep: standard, Riemann or delta-SPH pressure and viscosity,
  transport-velocity background and stress terms, artificial viscosity and
  heat conduction.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Dict, Union

import numpy as np
import torch

from . import _lib
from .eos import RIEMANNEoS, TaitEoS
from .jax_md.partition import NeighborList
from .jax_md.space import PeriodicDisplacement
from .kernel import KERNELS

_f = np.float32
EPS = float(np.finfo(float).eps)  # solver.py:21

STATE_KEYS = ("r", "tag", "u", "v", "drhodt", "dudt", "dvdt", "rho", "p", "mass", "eta",
              "dTdt", "T", "kappa", "Cp", "nw")


class GravityField:
    """A constant body force ``g`` as a ``g_ext_fn`` (e.g. ``db.py:294-297``).

    Called with positions it returns ``g`` broadcast to (N, dim); the fused
    step recognises it and applies ``g`` inside the force sweep."""

    def __init__(self, g):
        self.g = np.asarray(g, dtype=np.float32).reshape(-1)

    def __call__(self, r):
        g = torch.as_tensor(self.g, device=r.device, dtype=r.dtype)
        return g.expand(r.shape[0], -1).contiguous()


def zero_field(dim):
    """``_external_acceleration_fn`` returning zeros (tgv.py:53-54, ldc.py:127-128)."""
    return GravityField(np.zeros(dim, np.float32))


class WCSPH:
    """Weakly compressible SPH solver with transport velocity formulation."""

    def __init__(self, displacement_fn: Callable, eos: Union[TaitEoS, RIEMANNEoS],
                 g_ext_fn: Callable, dx: float, dim: int, dt: float, c_ref: float,
                 eta_limiter: float = 3, diff_delta: float = 0.02, diff_alpha: float = 0.1,
                 solver: str = "SPH", kernel: str = "QSK", h_fac: float = 1.0,
                 is_bc_trick: bool = False, is_rho_evol: bool = False,
                 artificial_alpha: float = 0.0, is_free_slip: bool = False,
                 is_rho_renorm: bool = False, is_heat_conduction: bool = False):
        if solver not in ("SPH", "RIE", "DELTA"):
            raise ValueError(f"unknown solver {solver!r} (SPH, RIE or DELTA)")
        if kernel not in KERNELS:
            raise ValueError(f"unknown kernel {kernel!r}")
        if not isinstance(displacement_fn, PeriodicDisplacement):
            raise NotImplementedError("WCSPH needs the displacement of space.periodic")
        self.displacement_fn = displacement_fn
        self.solver = solver
        self.g_ext_fn = g_ext_fn
        self.is_bc_trick = is_bc_trick
        self.is_rho_evol = is_rho_evol
        self.is_rho_renorm = is_rho_renorm
        self.dt = dt
        self.eos = eos
        self.c_ref = c_ref
        self.diff_delta = diff_delta
        self.diff_alpha = diff_alpha
        self.eta_limiter = eta_limiter
        self.artificial_alpha = artificial_alpha
        self.is_free_slip = is_free_slip
        self.is_heat_conduction = is_heat_conduction
        self.dx = dx
        self.dim = dim
        self.support = h_fac * dx          # solver.py:688, 693
        self._kernel_fn = KERNELS[kernel](h=h_fac * dx, dim=dim)

    # ---------------------------------------------------------------- params
    @property
    def g_const(self):
        """Constant body force when g_ext_fn is a GravityField, else None."""
        if isinstance(self.g_ext_fn, GravityField):
            return self.g_ext_fn.g
        return None

    def params(self, tvf_stress: bool = True) -> _lib.SolverParams:
        p = _lib.SolverParams()
        k = self._kernel_fn
        p.dim = self.dim
        p.kernel = k.kind
        p.solver = {"SPH": 0, "RIE": 1, "DELTA": 2}[self.solver]
        p.rho_renorm = 1 if self.is_rho_renorm else 0
        p.diff_delta = float(_f(self.diff_delta))
        p.diff_alpha = float(_f(self.diff_alpha))
        p.support = float(_f(self.support))
        p.density_evolution = 1 if self.is_rho_evol else 0
        p.bc_trick = 1 if self.is_bc_trick else 0
        p.free_slip = 1 if self.is_free_slip else 0
        p.heat_conduction = 1 if self.is_heat_conduction else 0
        p.one_over_h = k.one_over_h_f32
        p.sigma = k.sigma_f32
        p.dt = float(_f(self.dt))
        p.c_ref = float(_f(self.c_ref))
        p.eta_limiter = float(_f(self.eta_limiter))
        p.artificial_alpha = float(_f(self.artificial_alpha))
        p.av_h = float(_f(self.dx))
        e = self.eos
        if isinstance(e, RIEMANNEoS):
            p.eos_kind = 1
            p.rho_ref, p.p_bg = float(_f(e.rho_ref)), float(_f(e.p_bg))
            p.riemann_k = float(_f(100 * e.u_ref ** 2))
            p.p_ref, p.gamma, p.inv_gamma = 1.0, 1.0, 1.0
        elif isinstance(e, TaitEoS):
            p.eos_kind = 0
            p.p_ref, p.rho_ref, p.p_bg = float(_f(e.p_ref)), float(_f(e.rho_ref)), float(_f(e.p_bg))
            p.gamma, p.inv_gamma = float(_f(e.gamma)), float(_f(1.0 / e.gamma))
            p.riemann_k = 1.0
        else:
            raise NotImplementedError("eos must be TaitEoS or RIEMANNEoS")
        side = self.displacement_fn.side
        for d in range(3):
            p.box[d] = float(side[d]) if d < len(side) else 0.0
        g = self.g_const
        for d in range(3):
            p.g_const[d] = float(g[d]) if (g is not None and d < len(g)) else 0.0
        p.tvf_stress = 1 if tvf_stress else 0
        return p

    # -------------------------------------------------------------- forward
    def state_in(self, state: Dict[str, torch.Tensor]):
        """Normalised device tensors for the C-ABI (float32 contiguous, int32 tag)."""
        dev = _lib.device()
        s = {k: _lib.as_f32(state[k], dev) for k in STATE_KEYS if k != "tag"}
        s["tag"] = _lib.as_i32(state["tag"], dev)
        for k in ("r", "u", "v", "dudt", "dvdt", "nw"):
            if s[k].dim() != 2 or s[k].shape[1] != self.dim:
                raise ValueError(f"state[{k!r}] must be (N, {self.dim})")
        return s

    def c_state_in(self, s, g_ext=None) -> _lib.StateIn:
        c = _lib.StateIn()
        for k in _lib._IN_FIELDS:
            if k == "g_ext":
                setattr(c, k, g_ext.data_ptr() if g_ext is not None else None)
            else:
                setattr(c, k, s[k].data_ptr())
        return c

    def g_ext_tensor(self, r):
        if self.g_const is not None:
            return None
        g = self.g_ext_fn(r)
        return _lib.as_f32(g)

    def new_outputs(self, s):
        out = {k: torch.empty_like(s[k]) for k in ("u", "v", "dudt", "dvdt", "rho", "p",
                                                   "drhodt")}
        if self.is_heat_conduction:
            out["T"] = torch.empty_like(s["T"])
            out["dTdt"] = torch.empty_like(s["dTdt"])
        return out

    @staticmethod
    def c_state_out(out) -> _lib.StateOut:
        c = _lib.StateOut()
        for k in _lib._OUT_FIELDS:
            setattr(c, k, out[k].data_ptr() if k in out else None)
        return c

    @staticmethod
    def assemble(s, out):
        """solver.py:930-947: the 16-key output dict."""
        res = {k: s[k] for k in STATE_KEYS}
        res.update(out)
        return res

    def forward_wrapper(self):
        """Wrapper of update step of SPH."""

        def forward(state, neighbors):
            if not isinstance(neighbors, NeighborList):
                raise TypeError("neighbors must come from jax_sph_amd.partition.neighbor_list")
            ws = neighbors.workspace
            s = self.state_in(state)
            if s["r"].shape[0] != ws.n:
                raise ValueError("state size does not match the neighbour list")
            neighbors._check_live()
            if neighbors.bank == _lib.SKIN_BANK and \
                    s["r"].data_ptr() != neighbors.reference_position.data_ptr():
                # a skin list is filtered at the positions it belongs to; other
                # positions get an exact list of those pairs
                neighbors = neighbors._search.exact_at_reference(neighbors)
            out = self.new_outputs(s)
            g = self.g_ext_tensor(s["r"])
            prm = self.params(tvf_stress=True)
            _lib.check(_lib.lib().sph_forward(ws.handle, neighbors.bank, ctypes.byref(prm),
                                              ctypes.byref(self.c_state_in(s, g)),
                                              ctypes.byref(self.c_state_out(out)),
                                              _lib.stream()), "sph_forward")
            return self.assemble(s, out)

        forward._sph_solver = self
        return forward

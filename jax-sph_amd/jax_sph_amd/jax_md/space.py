"""Periodic space (drop-in for ``jax_sph/jax_md/space.py:170-286``).

``periodic(side)`` returns callable objects that carry ``side`` so the fused
HIP step can recognise them and apply the same geometry natively.  Called
directly they evaluate the reference formulas with torch on the tensor's device.
"""
from __future__ import annotations

import numpy as np
import torch


def _t(x, like: torch.Tensor):
    return torch.as_tensor(np.asarray(x, dtype=np.float32), device=like.device)


def _jmod(x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """jnp.mod: C fmod, then shift towards the divisor's sign."""
    r = torch.fmod(x, y)
    adj = (r != 0) & ((r < 0) != (y < 0))
    return torch.where(adj, r + y, r)


def periodic_displacement(side, dR: torch.Tensor) -> torch.Tensor:
    """space.py:170-181: mod(dR + side/2, side) - side/2."""
    s = _t(side, dR)
    half = s * 0.5
    return _jmod(dR + half, s) - half


def periodic_shift(side, R: torch.Tensor, dR: torch.Tensor) -> torch.Tensor:
    """space.py:207-209."""
    return _jmod(R + dR, _t(side, R))


def pairwise_displacement(Ra, Rb):
    return Ra - Rb


def square_distance(dR: torch.Tensor) -> torch.Tensor:
    """space.py:184-192 (x^2 + y^2 [+ z^2], in order)."""
    acc = dR[..., 0] * dR[..., 0]
    for k in range(1, dR.shape[-1]):
        acc = acc + dR[..., k] * dR[..., k]
    return acc


def distance(dR: torch.Tensor) -> torch.Tensor:
    """space.py:195-204, 0 at 0."""
    d2 = square_distance(dR)
    return torch.where(d2 > 0, torch.sqrt(torch.where(d2 > 0, d2, torch.zeros_like(d2))),
                       torch.zeros_like(d2))


class PeriodicDisplacement:
    """``displacement_fn`` of ``space.periodic`` with its box attached."""

    def __init__(self, side):
        self.side = np.asarray(side, dtype=np.float32).reshape(-1)

    def __call__(self, Ra, Rb, perturbation=None, **kwargs):
        if "box" in kwargs:
            raise ValueError("`space.periodic` does not accept a box argument.")
        dR = periodic_displacement(self.side, pairwise_displacement(Ra, Rb))
        if perturbation is not None:
            dR = dR * perturbation
        return dR


class PeriodicShift:
    """``shift_fn`` of ``space.periodic`` (wrapped into [0, side))."""

    def __init__(self, side, wrapped=True):
        self.side = np.asarray(side, dtype=np.float32).reshape(-1)
        self.wrapped = wrapped

    def __call__(self, R, dR, **kwargs):
        if "box" in kwargs:
            raise ValueError("`space.periodic` does not accept a box argument.")
        if not self.wrapped:
            return R + dR
        return periodic_shift(self.side, R, dR)


def periodic(side, wrapped: bool = True):
    """space.py:232-286: (displacement_fn, shift_fn) on a periodic hypercube."""
    return PeriodicDisplacement(side), PeriodicShift(side, wrapped)

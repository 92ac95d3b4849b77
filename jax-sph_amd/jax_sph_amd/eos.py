"""Equations of state (drop-in for ``jax_sph/eos.py``).

The classes carry the constants the fused HIP sweeps use; ``p_fn``/``rho_fn``
called directly evaluate the same float32 formulas on device tensors.
"""
from __future__ import annotations

import numpy as np
import torch

_f = np.float32


class BaseEoS:
    """eos.py:6-17."""

    kind = None

    def p_fn(self, rho):
        raise NotImplementedError

    def rho_fn(self, p):
        raise NotImplementedError


class TaitEoS(BaseEoS):
    """eos.py:20-38 (Adami et al. 2012)."""

    kind = 0

    def __init__(self, p_ref, rho_ref, p_background, gamma):
        self.p_ref = p_ref
        self.rho_ref = rho_ref
        self.p_bg = p_background
        self.gamma = gamma

    def p_fn(self, rho):
        x = rho / float(_f(self.rho_ref))
        if self.gamma != 1.0:
            x = torch.pow(x, float(_f(self.gamma))) if isinstance(x, torch.Tensor) else x ** self.gamma
        return float(_f(self.p_ref)) * (x - 1.0) + float(_f(self.p_bg))

    def rho_fn(self, p):
        p_temp = p + float(_f(self.p_ref)) - float(_f(self.p_bg))
        x = p_temp / float(_f(self.p_ref))
        if self.gamma != 1.0:
            e = float(_f(1.0 / self.gamma))
            x = torch.pow(x, e) if isinstance(x, torch.Tensor) else x ** e
        return float(_f(self.rho_ref)) * x


class RIEMANNEoS(BaseEoS):
    """eos.py:41-57 (Zhang, Hu, Adams 2017)."""

    kind = 1

    def __init__(self, rho_ref, p_background, u_ref):
        self.rho_ref = rho_ref
        self.u_ref = u_ref
        self.p_bg = p_background

    def p_fn(self, rho):
        return float(_f(100 * self.u_ref ** 2)) * (rho - float(_f(self.rho_ref))) + float(
            _f(self.p_bg))

    def rho_fn(self, p):
        return (p - float(_f(self.p_bg))) / float(_f(100 * self.u_ref ** 2)) + float(
            _f(self.rho_ref))

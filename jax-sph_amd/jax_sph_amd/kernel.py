"""SPH interpolation kernels (drop-in for ``jax_sph/kernel.py``).

Each class carries the reference's constants (``_one_over_h``, ``_sigma``,
``cutoff``); ``w`` and ``grad_w`` evaluate on the HIP device through
``sph_kernel_eval``.  ``grad_w`` is the closed-form derivative, equal to the
reference's ``jax.grad(w)`` wherever W is differentiable.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib


class BaseKernel:
    """kernel.py:9-23."""

    name = None
    _normalized_cutoff = None

    def __init__(self, h: float, dim: int = 3):
        self._one_over_h = 1.0 / h
        self.dim = dim
        self.h = h
        self.cutoff = self._normalized_cutoff * h
        self._sigma = self._sigma_of(dim)

    def _sigma_of(self, dim):  # pragma: no cover - overridden
        raise NotImplementedError

    @property
    def kind(self) -> int:
        return _lib.KERNEL_IDS[self.name]

    @property
    def sigma_f32(self) -> float:
        return float(np.float32(self._sigma))

    @property
    def one_over_h_f32(self) -> float:
        return float(np.float32(self._one_over_h))

    def _eval(self, r, want_w, want_dw):
        t = _lib.as_f32(r)
        flat = t.reshape(-1)
        w = torch.empty_like(flat) if want_w else None
        dw = torch.empty_like(flat) if want_dw else None
        _lib.check(_lib.lib().sph_kernel_eval(self.kind, self.dim, self.one_over_h_f32,
                                              self.sigma_f32, _lib.ptr(flat), flat.numel(),
                                              _lib.ptr(w), _lib.ptr(dw), _lib.stream()),
                   "sph_kernel_eval")
        shape = t.shape
        return (w.reshape(shape) if w is not None else None,
                dw.reshape(shape) if dw is not None else None)

    def w(self, r):
        """Kernel value at radial distance r."""
        return self._eval(r, True, False)[0]

    def grad_w(self, r):
        """dW/dr at radial distance r."""
        return self._eval(r, False, True)[1]


class CubicKernel(BaseKernel):
    """kernel.py:26-48."""

    name, _normalized_cutoff = "CSK", 2.0

    def _sigma_of(self, dim):
        ih = self._one_over_h
        return {1: 2.0 / 3.0 * ih, 2: 10.0 / 7.0 / math.pi * ih ** 2, 3: 1.0 / math.pi * ih ** 3}[dim]


class QuinticKernel(BaseKernel):
    """kernel.py:51-72 (the 3D constant 3/(359 pi) is kept as in the reference)."""

    name, _normalized_cutoff = "QSK", 3.0

    def _sigma_of(self, dim):
        ih = self._one_over_h
        return {1: 1.0 / 120.0 * ih, 2: 7.0 / 478.0 / math.pi * ih ** 2,
                3: 3.0 / 359.0 / math.pi * ih ** 3}[dim]


class WendlandC2Kernel(BaseKernel):
    """kernel.py:75-103."""

    name, _normalized_cutoff = "WC2K", 2.0

    def _sigma_of(self, dim):
        ih = self._one_over_h
        return {1: 5.0 / 8.0 * ih, 2: 7.0 / 4.0 / math.pi * ih ** 2,
                3: 21.0 / 16.0 / math.pi * ih ** 3}[dim]


class WendlandC4Kernel(BaseKernel):
    """kernel.py:106-134."""

    name, _normalized_cutoff = "WC4K", 2.0

    def _sigma_of(self, dim):
        ih = self._one_over_h
        return {1: 3.0 / 4.0 * ih, 2: 9.0 / 4.0 / math.pi * ih ** 2,
                3: 495.0 / 256.0 / math.pi * ih ** 3}[dim]


class WendlandC6Kernel(BaseKernel):
    """kernel.py:137-165."""

    name, _normalized_cutoff = "WC6K", 2.0

    def _sigma_of(self, dim):
        ih = self._one_over_h
        return {1: 55.0 / 64.0 * ih, 2: 78.0 / 28.0 / math.pi * ih ** 2,
                3: 1365.0 / 512.0 / math.pi * ih ** 3}[dim]


class GaussianKernel(BaseKernel):
    """kernel.py:168-182."""

    name, _normalized_cutoff = "GK", 3.0

    def _sigma_of(self, dim):
        return 1.0 / math.pi ** (dim / 2) * self._one_over_h ** dim


class SuperGaussianKernel(BaseKernel):
    """kernel.py:185-201."""

    name, _normalized_cutoff = "SGK", 3.0

    def _sigma_of(self, dim):
        return 1.0 / math.pi ** (dim / 2) * self._one_over_h ** dim


KERNELS = {"CSK": CubicKernel, "QSK": QuinticKernel, "WC2K": WendlandC2Kernel,
           "WC4K": WendlandC4Kernel, "WC6K": WendlandC6Kernel, "GK": GaussianKernel,
           "SGK": SuperGaussianKernel}

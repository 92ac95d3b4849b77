"""MI355X-native SPH step with the JAX-SPH interface.

Drop-in modules (same names and call signatures as ``jax_sph``):
``partition``, ``kernel``, ``eos``, ``solver``, ``integrator``, ``utils``,
``jax_md.space``, ``jax_md.partition``; plus ``case_setup`` for the cases the
step is measured on.  All per-step work runs in ``libsph.so`` (HIP, gfx950).
"""
__version__ = "0.1.0"

from . import eos, jax_md, kernel, utils  # noqa: F401,E402
from . import partition, solver, integrator, case_setup  # noqa: F401,E402

"""Drop-in for ``jax_sph.jax_md``: the parts of JAX-MD the SPH step uses."""
from . import partition, space  # noqa: F401

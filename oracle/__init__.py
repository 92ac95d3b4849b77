"""CPU oracle for the SPH step: test infrastructure, never product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
path (``jax-sph_amd``) never imports it and has no CPU fallback.
"""

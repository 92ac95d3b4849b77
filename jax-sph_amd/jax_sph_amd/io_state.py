"""State files (stand-in for ``jax_sph/io_state.py``).

The reference writes one h5 dataset per state key (io_state.py:14-85) with
h5py, which this package does not depend on.  Here a state is written as an
``.npz`` archive with the same layout: one array per state key, float32
vectors and scalars and an int32 ``tag``.  ``read_state`` returns numpy arrays;
:meth:`case_setup.SimulationSetup.initialize` accepts such a file as
``cfg.case.state0_path`` (the relaxed positions of a relaxation run, as
case_setup.py:184-195 does with an h5 file).
"""
from __future__ import annotations

from typing import Dict

import numpy as np
import torch


def write_state(path: str, state: Dict) -> None:
    arrays = {}
    for k, v in state.items():
        a = v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)
        arrays[k] = a.astype(np.int32) if k == "tag" else a.astype(np.float32)
    np.savez(path, **arrays)


def read_state(path: str) -> Dict[str, np.ndarray]:
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}

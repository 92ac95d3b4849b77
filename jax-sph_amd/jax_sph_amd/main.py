"""Command-line entry point (stand-in for the reference's ``main.py``)::

    python -m jax_sph_amd.main config=cases/tgv.yaml case.dx=0.02 solver.name=RIE

Arguments are ``key=value`` with dotted config keys. Values are parsed as YAML
(numbers, booleans, lists, ``null``), as ``OmegaConf.from_cli`` does.
``config`` names the YAML case file, whose ``extends`` chain ends at
``JAX_SPH_DEFAULTS`` (main.py:20-47). The run is ``simulate(cfg)`` with the
reference's trajectory writer (io_state.py:14-68):
* states every ``io.write_every`` steps, as ``traj_<step>`` files in a run
  directory under ``io.data_path``;
* for a relaxation (``case.mode=rlx``), the last state, as
  ``<case>_<dim>_<dx>_<seed>``;
* ``config.yaml`` of the run beside them.

Files are ``.npz`` (:func:`io_state.write_state`) where the reference writes
``.h5``; ``vtk`` output is not provided.
"""
from __future__ import annotations

import os
import sys
import time
from typing import Dict, List

import yaml

from .case_setup import load_embedded_configs
from .io_state import write_state

_FORMATS = ("h5", "npz")


def parse_cli(argv: List[str]) -> Dict:
    """``key.sub=value`` arguments as a nested dict of YAML-parsed values."""
    out: Dict = {}
    for arg in argv:
        if "=" not in arg:
            raise ValueError(f"arguments are key=value, got {arg!r}")
        key, val = arg.split("=", 1)
        node = out
        parts = key.split(".")
        for p in parts[:-1]:
            node = node.setdefault(p, {})
            if not isinstance(node, dict):
                raise ValueError(f"{key!r} sets a key below a value")
        node[parts[-1]] = yaml.safe_load(val) if val != "" else None
    return out


def _plain(x):
    if isinstance(x, dict):
        return {k: _plain(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_plain(v) for v in x]
    if hasattr(x, "tolist"):
        return x.tolist()
    return x


class TrajectoryWriter:
    """io_state.py:14-68: ``io_setup`` once the case is set up (``bind``),
    then ``write_state(step, state)`` for each step of the loop."""

    def __init__(self):
        self.cfg = None
        self.dir = None
        self.written: List[str] = []

    def bind(self, cfg) -> None:
        self.cfg = cfg
        d = cfg.io.data_path
        if len(cfg.io.write_type) > 0 and cfg.case.mode == "sim":
            case = os.path.splitext(os.path.basename(cfg.case.source or "case"))[0].upper()
            d = os.path.join(d, f"{cfg.case.dim}D_{case}_{cfg.solver.name}_{cfg.seed}_"
                             + time.strftime("%Y%m%d-%H%M%S"))
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "config.yaml"), "w") as f:
            yaml.safe_dump(_plain(cfg), f)
        self.dir = d

    def __call__(self, step: int, state: Dict) -> None:
        cfg = self.cfg
        if not any(t in cfg.io.write_type for t in _FORMATS):
            return
        step_max = cfg.solver.sequence_length
        normal = cfg.case.mode == "sim" and step >= 0 and step % cfg.io.write_every == 0
        relax = cfg.case.mode == "rlx" and step == step_max - 1
        if not (normal or relax):
            return
        if cfg.case.mode == "rlx":
            case = os.path.splitext(os.path.basename(cfg.case.source or "case"))[0]
            name = "_".join([case, str(cfg.case.dim), str(cfg.case.dx), str(cfg.seed)])
        else:
            name = "traj_" + str(step).zfill(len(str(step_max)))
        path = os.path.join(self.dir, name + ".npz")
        write_state(path, state)
        self.written.append(path)


def main(argv: List[str] | None = None) -> Dict:
    from .simulate import simulate

    cli = parse_cli(sys.argv[1:] if argv is None else argv)
    if "config" not in cli:
        raise ValueError("A configuration file must be specified.")
    path = cli.pop("config")
    cfg = load_embedded_configs(path, cli)
    writer = TrajectoryWriter()
    res = simulate(cfg, write_fn=writer)
    res["written"] = writer.written
    res["dir"] = writer.dir
    return res


if __name__ == "__main__":
    main()

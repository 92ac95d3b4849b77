"""The command-line run (main.py of the reference): a YAML case file, key=value
overrides, and the trajectory files the reference's writer names."""
import os
import tempfile

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

js = pytest.importorskip("jax_sph_amd")
from jax_sph_amd import case_setup as cs  # noqa: E402
from jax_sph_amd import integrator, partition  # noqa: E402
from jax_sph_amd import main as cli  # noqa: E402
from jax_sph_amd.io_state import read_state  # noqa: E402


def test_main_writes_the_reference_trajectory():
    with tempfile.TemporaryDirectory() as d:
        case = os.path.join(d, "tgv.yaml")
        with open(case, "w") as f:
            f.write("extends: JAX_SPH_DEFAULTS\ncase:\n  source: tgv.py\n  dim: 2\n"
                    "  dx: 0.05\nsolver:\n  t_end: 5.0\n")
        res = cli.main([f"config={case}", "solver.dt=0.0004", "solver.t_end=0.004",
                        "io.write_type=[h5]", "io.write_every=5", f"io.data_path={d}"])
        names = [os.path.basename(p) for p in res["written"]]
        assert names == ["traj_00.npz", "traj_05.npz", "traj_10.npz"]
        assert os.path.exists(os.path.join(res["dir"], "config.yaml"))
        t0 = read_state(res["written"][0])
        assert set(t0) >= {"r", "u", "v", "rho", "p", "tag", "dudt", "dvdt"}
    # traj_00 is the state after one step (simulate.py:114-117)
    cfg = cs.load_config("tgv", {"case": {"dim": 2, "dx": 0.05}, "solver": {"dt": 0.0004}})
    (cfg, box, state, g, bc, nw, eos, key, disp, shift) = cs.load_case("tgv")(cfg).initialize()
    solver = cs.make_solver(cfg, disp, eos, g)
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff, mask_self=False,
                                 format=partition.Sparse)
    adv = integrator.si_euler(cfg.solver.tvf, solver.forward_wrapper(), shift, bc, nw)
    st, nb = adv(0.0004, state, fn.allocate(state["r"]))
    np.testing.assert_array_equal(t0["r"], st["r"].cpu().numpy())
    np.testing.assert_array_equal(t0["u"], st["u"].cpu().numpy())

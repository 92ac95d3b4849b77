"""HIP-graph replay of fused steps equals the eager steps bit for bit, over
several replays with eager work and synchronisations in between."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

js = pytest.importorskip("jax_sph_amd")
from jax_sph_amd.graph import StepGraph  # noqa: E402
from tests.test_step_gpu import _make  # noqa: E402

KEYS = ("r", "u", "v", "dudt", "dvdt", "rho", "p", "drhodt")


@pytest.mark.parametrize("skin", ["auto", 0.0])
@pytest.mark.parametrize("case,ov", [
    ("tgv", {"case": {"dim": 2, "dx": 0.02}}),
    ("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / 16}, "solver": {"tvf": 1.0}}),
    ("ldc", {"case": {"dx": 0.04}}),
])
def test_graph_replays_equal_eager_steps(case, ov, skin):
    cfg, state, nb, adv, *_ = _make(case, ov, skin=skin)
    dt = cfg.solver.dt
    st = state
    for _ in range(12):
        st, nb = adv(dt, st, nb)
    eager = {k: st[k].clone() for k in KEYS}
    eager_bits = int(nb.info[0])

    _, state2, nb2, adv2, *_ = _make(case, ov, skin=skin)
    g = StepGraph(adv2, state2, nb2, dt, steps_per_replay=3)
    side = torch.zeros(1000, device="cuda")
    for k in range(4):
        g.replay()
        side += 1.0                      # eager work between replays
        torch.cuda.synchronize()
    assert g.replays == 4
    for k in KEYS:
        assert torch.equal(g.state[k], eager[k]), k
    assert int(g._info0[0]) == eager_bits
    assert not g.did_buffer_overflow

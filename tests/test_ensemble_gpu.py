"""GPU parity of the ensemble step: every member of a batched step equals its
own fused step bitwise (same kernels, same per-member arguments), under every
block layout, eager or graph-replayed, and after autotuning; one member is
also checked against the float32/float64 oracles directly."""
import numpy as np
import pytest
import torch

from oracle import nl_oracle as nlo
from oracle import sph_oracle as so

pytestmark = pytest.mark.gpu

js = pytest.importorskip("jax_sph_amd")
from jax_sph_amd import case_setup as cs  # noqa: E402
from jax_sph_amd import integrator, partition  # noqa: E402
from jax_sph_amd.ensemble import Ensemble  # noqa: E402
from jax_sph_amd.jax_md.partition import Sparse  # noqa: E402
from tests import gpu_helpers as G  # noqa: E402

FIELDS = ("r", "u", "v", "dudt", "dvdt", "rho", "p", "drhodt")

# mixed plans (four groups of equal sweeps): summation without walls, with
# walls, with the transport-velocity stress; Riemann continuity with walls and
# gravity on the Wendland kernel
MIXED = [
    ("tgv", {"case": {"dim": 2, "dx": 0.05}}),
    ("ldc", {"case": {"dx": 0.05}, "solver": {"dt": None}}),
    ("tgv", {"case": {"dim": 2, "dx": 0.04}}),
    ("db", {"case": {"dx": 0.04}, "solver": {"name": "RIE"}, "kernel": {"name": "WC2K"}}),
    ("ldc", {"case": {"dx": 0.04}, "solver": {"dt": None}}),
    ("tgv", {"case": {"dim": 2, "dx": 0.025}, "solver": {"tvf": 1.0}}),
]


def _cfgs(spec):
    return [cs.load_config(name, ov) for name, ov in spec]


def _single(cfg):
    """The member as its own simulation: case, list, fused si_euler."""
    Case = cs.load_case(cfg.case.source or cfg.config)
    (cfg, box, state, g, bc, nw_fn, eos, key, disp, shift) = Case(cfg).initialize()
    solver = cs.make_solver(cfg, disp, eos, g)
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False, format=Sparse)
    nb = fn.allocate(state["r"])
    adv = integrator.si_euler(cfg.solver.tvf, solver.forward_wrapper(), shift, bc, nw_fn)
    return cfg, box, state, nb, adv, solver, bc


def _run_single(cfg, steps):
    cfg, box, state, nb, adv, *_ = _single(cfg)
    st = state
    for _ in range(steps):
        st, nb = adv(cfg.solver.dt, st, nb)
    return st, nb


def _assert_equal(ens, k, st, nb, label):
    got = ens.state(k)
    for f in FIELDS:
        a, b = got[f].cpu().numpy(), st[f].cpu().numpy()
        assert np.array_equal(a, b), (label, k, f, np.abs(a - b).max())
    ia, ib = ens.error_info(k).cpu().numpy(), nb.info.cpu().numpy()
    # error bits, occupancy, largest cell, largest list
    assert np.array_equal(ia[[0, 1, 2, 3]], ib[[0, 1, 2, 3]]), (label, k, ia, ib)


def test_members_equal_their_own_fused_steps():
    cfgs = _cfgs(MIXED)
    ens = Ensemble.from_configs(cfgs)
    assert ens.info()["groups"] == 4
    steps = 7
    ens.step(steps)
    torch.cuda.synchronize()
    assert not ens.did_buffer_overflow
    for k, cfg in enumerate(_cfgs(MIXED)):
        st, nb = _run_single(cfg, steps)
        _assert_equal(ens, k, st, nb, "eager")


@pytest.mark.parametrize("order,xcd", [(0, False), (1, True), (2, True), (2, False)])
def test_layouts_and_graph_replay_agree(order, xcd):
    spec = MIXED[:3] + MIXED[4:]
    ref = Ensemble.from_configs(_cfgs(spec))
    ref.step(12)
    ens = Ensemble.from_configs(_cfgs(spec), order=order, xcd=xcd)
    ens.step(2)          # eager steps first, then graph replays continue them
    ens.capture(4)
    ens.replay(2)
    ens.step(2)
    torch.cuda.synchronize()
    assert ens.steps == 12
    for k in range(len(spec)):
        for f in FIELDS:
            a, b = ens.state(k)[f].cpu().numpy(), ref.state(k)[f].cpu().numpy()
            assert np.array_equal(a, b), (order, xcd, k, f)
        assert torch.equal(ens.error_info(k), ref.error_info(k))


def test_autotune_keeps_the_trajectory():
    spec = MIXED[:3]
    ref = Ensemble.from_configs(_cfgs(spec))
    ref.step(6)
    ens = Ensemble.from_configs(_cfgs(spec))
    res = ens.autotune(steps=4, reps=1)
    assert len(res["ms_per_step"]) == 5 and ens.steps == 0
    ens.step(6)
    torch.cuda.synchronize()
    for k in range(len(spec)):
        for f in FIELDS:
            assert torch.equal(ens.state(k)[f], ref.state(k)[f]), (k, f)


def test_3d_members_equal_their_own_fused_steps():
    """A 3D ensemble (two Taylor-Green cubes with the transport-velocity
    formulation, 16^3 and 17^3 particles; 17^3 is not a multiple of the
    256-particle integrate blocks): every member equals its own fused steps
    bitwise."""
    spec = [("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / 16}, "solver": {"tvf": 1.0}}),
            ("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / 17}, "solver": {"tvf": 1.0}})]
    ens = Ensemble.from_configs(_cfgs(spec))
    steps = 6
    ens.step(steps)
    torch.cuda.synchronize()
    assert not ens.did_buffer_overflow
    for k, cfg in enumerate(_cfgs(spec)):
        st, nb = _run_single(cfg, steps)
        _assert_equal(ens, k, st, nb, "3d")


def test_tuned_ensemble_members_equal_their_own_steps():
    """Ensemble.tuned_from_configs searches each group's Verlet-skin width
    and the block layout; the members of the ensemble it returns equal their
    own fused steps with the skin chosen for them, bitwise."""
    from jax_sph_amd.ensemble import setup_member
    spec = MIXED[:3]
    ens, rep = Ensemble.tuned_from_configs(_cfgs(spec), factors=(0.05, 0.15), steps=4, reps=1)
    assert len(rep["skin_factor_per_group"]) == len(rep["groups"]) == 2
    assert len(rep["trials_ms_per_step"]) >= 2
    assert ens.steps == 0
    steps = 6
    ens.step(steps)
    torch.cuda.synchronize()
    for k, (cfg, f) in enumerate(zip(_cfgs(spec), ens.skin_factors)):
        assert f in (0.05, 0.15)
        adv, st, nb, dt = setup_member(cfg, skin_factor=f)
        for _ in range(steps):
            st, nb = adv(dt, st, nb)
        _assert_equal(ens, k, st, nb, "tuned")


def test_member_matches_oracle():
    """A member after 5 batched steps against the float32/float64 oracles and
    the oracle's exact pair set at its new positions."""
    cfg = cs.load_config("ldc", {"case": {"dx": 0.04}, "solver": {"dt": None}})
    spec = [("tgv", {"case": {"dim": 2, "dx": 0.05}}),
            ("ldc", {"case": {"dx": 0.04}, "solver": {"dt": None}})]
    ens = Ensemble.from_configs(_cfgs(spec))
    ens.step(5)
    torch.cuda.synchronize()
    cfg, box, state, nb, adv, solver, bc = _single(cfg)
    ns = nlo.NeighborSearch(box, solver._kernel_fn.cutoff, mask_self=False)
    ns64 = nlo.NeighborSearch(box, solver._kernel_fn.cutoff, mask_self=False, dtype=np.float64)
    bcf = G.oracle_bc(bc, 2)
    adv32 = so.si_euler(cfg.solver.tvf, G.oracle_solver(solver, box).forward, box, bcf,
                        ns.update)
    adv64 = so.si_euler(cfg.solver.tvf, G.oracle_solver64(solver, box).forward, box, bcf,
                        ns64.update, np.float64)
    st32, st64 = G.to_np(state), G.to_f64(G.to_np(state))
    nb32, nb64 = ns.allocate(st32["r"]), ns64.allocate(st64["r"])
    for _ in range(5):
        st32, nb32 = adv32(cfg.solver.dt, st32, nb32)
        st64, nb64 = adv64(cfg.solver.dt, st64, nb64)
    got = ens.state(1)
    G.assert_within_noise_floor(got, st32, st64, 1e-5, fields=("r", "u", "v", "rho"),
                                label="ensemble ldc")
    r = got["r"].cpu().numpy()
    assert nlo.sorted_pairs(ns.allocate(r).idx, len(r)).shape[1] == int(ens.error_info(1)[1])


def test_rejects_what_it_cannot_batch():
    cfg = cs.load_config("tgv", {"case": {"dim": 2, "dx": 0.05}})
    Case = cs.load_case("tgv")
    (cfg, box, state, g, bc, nw_fn, eos, key, disp, shift) = Case(cfg).initialize()
    solver = cs.make_solver(cfg, disp, eos, g)
    adv = integrator.si_euler(0.0, solver.forward_wrapper(), shift, bc, None)
    fn0 = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                  mask_self=False, format=Sparse, skin=0.0)
    with pytest.raises(ValueError, match="skin"):
        Ensemble([(adv, state, fn0.allocate(state["r"]), cfg.solver.dt)])
    c3 = cs.load_config("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / 16}})
    (c3, box3, st3, g3, bc3, _, eos3, _, disp3, shift3) = Case(c3).initialize()
    s3 = cs.make_solver(c3, disp3, eos3, g3)
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 mask_self=False, format=Sparse)
    fn3 = partition.neighbor_list(disp3, box3, r_cutoff=s3._kernel_fn.cutoff,
                                  mask_self=False, format=Sparse)
    adv3 = integrator.si_euler(0.0, s3.forward_wrapper(), shift3, bc3, None)
    with pytest.raises(ValueError, match="dimension"):
        Ensemble([(adv, state, fn.allocate(state["r"]), cfg.solver.dt),
                  (adv3, st3, fn3.allocate(st3["r"]), c3.solver.dt)])
    with pytest.raises(ValueError, match="even"):
        Ensemble([(adv, state, fn.allocate(state["r"]), cfg.solver.dt)]).capture(3)

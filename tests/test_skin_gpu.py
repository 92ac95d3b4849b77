"""Verlet-skin lists of the fused step.

The reference rebuilds the neighbour list on every si_euler step
(integrator.py:46 with dr_threshold = 0).  The skin lists rebuild only when a
particle has moved far enough, and filter every pair with the reference
predicate on every step; these tests check, step by step and across several
device-side rebuilds, that the pair set, the occupancy, the largest list of one
particle and the error bits equal the oracle's per-step rebuild, and that the
physics matches the oracle at the float32 noise floor.
"""
import numpy as np
import pytest
import torch

from oracle import nl_oracle as nlo

pytestmark = pytest.mark.gpu

js = pytest.importorskip("jax_sph_amd")
from tests import gpu_helpers as G  # noqa: E402
from tests.test_step_gpu import STATE, _make  # noqa: E402

TGV2 = {"case": {"dim": 2, "dx": 0.02}}
CASES = [  # case, overrides, skin, steps, rebuilds expected at least
    ("tgv", TGV2, "auto", 40, 3),
    ("tgv", TGV2, 0.0006, 12, 6),    # skin 1% of the cutoff: a rebuild nearly every step
    ("tgv", TGV2, 0.03, 40, 1),      # skin half the cutoff
    ("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / 16}, "solver": {"tvf": 1.0}}, 0.01, 16, 3),
    # boxes of a whole number of cutoffs (reference cells exactly the cutoff
    # wide: 10 cells of 0.1 in 2D, 6 of 2 pi / 6 in 3D): pairs in the band are
    # also tested for adjacent reference cells
    ("tgv", {"case": {"dim": 2, "dx": 1 / 30}}, "auto", 40, 2),
    ("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / 18}, "solver": {"tvf": 1.0}}, "auto", 24, 2),
    # 17^3 = 4,913 particles: not a multiple of the 256-thread blocks, so the
    # 3D integrate's partial last block (LDS-staged loads and stores) runs
    ("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / 17}, "solver": {"tvf": 1.0}}, 0.01, 16, 2),
    ("ldc", {"case": {"dx": 0.04}}, "auto", 30, 2),
    # dam break at rest: the first build only (the walls and the column barely move)
    ("db", {"case": {"dx": 0.04}, "solver": {"name": "RIE"}, "kernel": {"name": "WC2K"}},
     0.002, 30, 1),
]


@pytest.mark.parametrize("case,ov,skin,steps,min_rebuilds", CASES)
def test_skin_lists_equal_per_step_rebuild(case, ov, skin, steps, min_rebuilds):
    cfg, state, nb, adv, R, fn, *_ = _make(case, ov, skin=skin)
    assert nb.workspace.skin_on, "this geometry should take skin lists"
    r0 = G.to_np(state)["r"]
    ref_nb = R.ns.allocate(r0)
    assert nb.max_occupancy == ref_nb.max_occupancy
    st = state
    rebuilds = 0
    overflowed = False
    for k in range(steps):
        st, nb = adv(cfg.solver.dt, st, nb)
        assert nb.bank == 2
        info = nb.info.cpu().numpy()
        rebuilds += int(info[7] > 0)     # info[7] is set only by a rebuild
        r = st["r"].cpu().numpy()
        ref_nb = R.ns.update(r, ref_nb)
        n = len(r)
        assert int(info[0]) == int(ref_nb.error), f"step {k}: error bits"
        if int(ref_nb.error):
            # an overflow: the reference's pair set is then incomplete by
            # design; both sides re-allocate from these positions, as
            # simulate.py:119-130 does, and the run goes on
            nb = fn.allocate(st["r"])
            ref_nb = R.ns.allocate(r)
            overflowed = True
            continue
        got = nlo.sorted_pairs(nb.idx.cpu().numpy(), n)
        want = nlo.sorted_pairs(ref_nb.idx, n)
        assert got.shape == want.shape and np.array_equal(got, want), f"step {k}: pair sets"
        assert int(info[1]) == ref_nb.occupancy, f"step {k}: occupancy"
        per_sender = np.bincount(want[1], minlength=n)
        assert int(info[3]) == int(per_sender.max()), f"step {k}: largest list"
        assert int(info[0]) == int(ref_nb.error), f"step {k}: error bits"
        occ = nlo.count_cell_filling(r, R.ns.box, R.ns.cutoff, clamp=True).max()
        assert int(info[2]) == int(occ), f"step {k}: largest cell occupancy"
    assert rebuilds >= min_rebuilds, f"only {rebuilds} rebuilds in {steps} steps"
    if overflowed:
        return  # the physics of a step on an overflowed list is not the reference's
    ref, ref64 = R.run(state, steps), R.run(state, steps, f64=True)
    G.assert_within_noise_floor(st, ref, ref64, 1e-5, fields=STATE, label=f"{case} skin")


def test_skin_buffer_overflow_reallocates():
    """A skin list that outgrows its buffer reports NEIGHBOR_LIST_OVERFLOW; the
    re-allocation (simulate.py:119-130) sizes the buffer from the count seen."""
    cfg, state, nb, adv, R, fn, *_ = _make("tgv", TGV2, skin=0.006)
    fn.search._skin_width_override = 1
    try:
        nb = fn.allocate(state["r"])
    finally:
        fn.search._skin_width_override = None
    assert nb.workspace.skin_on
    st1, nb1 = adv(cfg.solver.dt, state, nb)
    assert bool(nb1.did_buffer_overflow)
    assert int(nb1.info[0]) & 1 and int(nb1.info[5]) > 0
    nb = fn.allocate(state["r"])
    st, nb = adv(cfg.solver.dt, state, nb)
    assert not bool(nb.did_buffer_overflow)
    ref, ref64 = R.run(state, 1), R.run(state, 1, f64=True)
    G.assert_within_noise_floor(st, ref, ref64, 1e-5, fields=STATE, label="skin realloc")


def test_forward_on_skin_list():
    """WCSPH.forward with a list from the fused step: at that step's positions it
    sweeps the skin lists; at other positions it uses an exact list of the same
    pairs, like forward(state, neighbors) with a stored idx."""
    cfg, state, nb, adv, R, fn, shift, solver, bc = _make("tgv", TGV2)
    st, nb = adv(cfg.solver.dt, state, nb)
    model = solver.forward_wrapper()
    out_skin = model(st, nb)
    np_st = G.to_np(st)
    ref_nl = nlo.NeighborSearch(np.ones(2), solver._kernel_fn.cutoff, mask_self=False).allocate(
        np_st["r"])
    ref = G.oracle_solver(solver, np.ones(2)).forward(np_st, ref_nl.idx)
    ref64 = G.oracle_solver64(solver, np.ones(2)).forward(G.to_f64(np_st), ref_nl.idx)
    G.assert_within_noise_floor(out_skin, ref, ref64, 1e-5, label="forward on skin list")
    # same values, another tensor: the exact list at the list's own positions
    st2 = dict(st)
    st2["r"] = st["r"].clone() + 0.0
    out2 = model(st2, nb)
    G.assert_within_noise_floor(out2, ref, ref64, 1e-5, label="forward via exact list")
    torch.cuda.synchronize()


def test_3d_integrate_partial_block_bitwise():
    """One fused step of a 3D case whose particle count is not a multiple of
    the integrate kernel's 256-particle blocks (17^3 = 4,913), from a state
    with nonzero dudt and dvdt: the kick, transport velocity and drift of the
    skin-list step (k_integrate_skin, staged through LDS) equal those of the
    exact-list step bit for bit, tail block included."""
    ov = {"case": {"dim": 3, "dx": 2 * np.pi / 17}, "solver": {"tvf": 1.0}}
    out = {}
    for skin in ("auto", 0.0):
        cfg, state, nb, adv, *_ = _make("tgv", ov, skin=skin)
        assert nb.workspace.skin_on == (skin == "auto")
        n = state["r"].shape[0]
        assert n == 4913 and n % 256
        gen = torch.Generator().manual_seed(3)
        st = dict(state)
        st["dudt"] = torch.randn(n, 3, generator=gen).cuda()
        st["dvdt"] = torch.randn(n, 3, generator=gen).cuda()
        st, nb = adv(cfg.solver.dt, st, nb)
        out[skin] = {k: st[k].cpu().numpy() for k in ("r", "u", "v")}
    for k in ("r", "u", "v"):
        assert np.array_equal(out["auto"][k], out[0.0][k]), k

"""Full-size runs of the benchmark configurations, checked through properties
that do not depend on size (the oracle is too slow at these sizes):

* 3D Taylor-Green vortex, 64^3 = 262,144 particles, tvf = 1, 500 steps:
  momentum balance of the right-hand side, kinetic-energy decay, finite
  fields, no overflow, and the final pair set against brute force;
* 2D dam break at dx = 0.005 (88,874 particles, Riemann SPH, Wendland C2)
  through the collapse with re-allocation on overflow: walls stay put, the
  front advances, fields stay finite; and oracle parity at dx = 0.04 once
  the column moves;
* the 64-member batched ensemble of 2D Taylor-Green vortices and lid-driven
  cavities (872,572 particles): the largest members equal their own fused
  steps bitwise.
"""
import numpy as np
import pytest
import torch

from oracle import nl_oracle as nlo

pytestmark = pytest.mark.gpu

js = pytest.importorskip("jax_sph_amd")
from jax_sph_amd import case_setup as cs  # noqa: E402
from jax_sph_amd import integrator, partition  # noqa: E402
from jax_sph_amd.utils import Tag  # noqa: E402
from tests import gpu_helpers as G  # noqa: E402


def _momentum_residual(mass, a):
    """max over components of |sum m a| / sum |m a| (float64 sums)."""
    ma = mass.double()[:, None] * a.double()
    return (ma.sum(0).abs() / ma.abs().sum(0).clamp_min(1e-300)).max()


def test_tgv3d_64cubed_500_steps():
    """The 3D benchmark configuration run for its 500 steps.  Every pair term
    of the standard + transport-velocity acceleration (solver.py:199-256) is
    antisymmetric in (i, j) once multiplied by m_i, and so is the
    background-pressure term, so sum_i m_i dudt_i and sum_i m_i dvdt_i vanish
    in a periodic box up to float32 rounding of the pair terms.  The bounds
    are relative to sum_i |m_i a_i|: measured 1.6e-6 (dudt) and 6.6e-4 (dvdt,
    whose per-particle sums cancel to a small remainder on the near-lattice, so
    the rounding of its ~113 pair terms is large next to it); a term that
    breaks the antisymmetry leaves a residual of the order of the field.
    Viscosity makes the kinetic energy of u decay over the run; from step to
    step it also trades with the acoustic energy of the weakly compressible
    fluid (measured rises up to 3e-4 per step)."""
    n_side = 64
    ov = {"case": {"dim": 3, "dx": 2 * np.pi / n_side, "viscosity": 1 / 1600},
          "solver": {"tvf": 1.0, "t_end": 1.0}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("tgv", ov)
    assert state["r"].shape[0] == n_side ** 3
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse)
    nb = fn.allocate(state["r"])
    adv = integrator.si_euler(cfg.solver.tvf, solver.forward_wrapper(), shift, bc, None)
    assert adv._sph_fusable(nb) is not None
    steps = 500
    mass = state["mass"]
    res_u = torch.zeros(steps, dtype=torch.float64, device="cuda")
    res_v = torch.zeros_like(res_u)
    ke = torch.zeros(steps + 1, dtype=torch.float64, device="cuda")
    ke[0] = 0.5 * (mass.double()[:, None] * state["u"].double() ** 2).sum()
    st = state
    for k in range(steps):
        st, nb = adv(cfg.solver.dt, st, nb)
        res_u[k] = _momentum_residual(mass, st["dudt"])
        res_v[k] = _momentum_residual(mass, st["dvdt"])
        ke[k + 1] = 0.5 * (mass.double()[:, None] * st["u"].double() ** 2).sum()
    torch.cuda.synchronize()
    assert not bool(nb.did_buffer_overflow)
    for f in ("r", "u", "v", "dudt", "dvdt", "rho", "p"):
        assert torch.isfinite(st[f]).all(), f
    ru, rv = float(res_u.max()), float(res_v.max())
    rise = float(((ke[1:] - ke[:-1]) / ke[:-1]).max())
    print(f"tgv3d 64^3: momentum residual dudt {ru:.2e} dvdt {rv:.2e}; "
          f"largest kinetic-energy rise per step {rise:.2e}; "
          f"E_kin {float(ke[0]):.6f} -> {float(ke[-1]):.6f}", flush=True)
    assert ru <= 1e-5 and rv <= 5e-3, (ru, rv)
    assert rise <= 1e-3, rise
    assert float(ke[-1]) < 0.95 * float(ke[0])
    # the pair set at the final positions: 512 sampled senders against brute force
    r = st["r"].cpu().numpy()
    idx = nb.idx.cpu().numpy()
    N = len(r)
    assert nb.occupancy == int((idx[0] < N).sum())
    t = np.float32
    cut = solver._kernel_fn.cutoff
    c2 = t(t(cut) * t(cut))
    side = np.asarray(box, np.float32)
    for s in np.random.default_rng(0).choice(N, 512, replace=False):
        d = nlo.square_distance(nlo.periodic_displacement(side, r[s][None, :] - r))
        want = np.sort(np.nonzero(d < c2)[0])
        got = np.sort(idx[0][idx[1] == s])
        np.testing.assert_array_equal(got, want)


def _dam_break(dx, dt=None):
    ov = {"case": {"dx": dx}, "solver": {"name": "RIE", "dt": dt}, "kernel": {"name": "WC2K"}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("db", ov)
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse)
    adv = integrator.si_euler(cfg.solver.tvf, solver.forward_wrapper(), shift, bc, None)
    return cfg, box, state, fn, adv, solver, bc


def test_dam_break_full_size_through_the_collapse():
    """Dam break at dx = 0.005: 88,874 particles (80,000 fluid), the CFL time
    step (8.0e-5; the case file's 3e-4 exceeds it at this resolution), 6,000
    steps (t = 0.48: the front leaves the column and keeps advancing).  The
    host loop re-allocates on overflow as simulate.py:119-130 does."""
    cfg, box, state, fn, adv, solver, bc = _dam_break(0.005)
    n = state["r"].shape[0]
    assert n == 88874
    tag = state["tag"]
    fluid = tag == int(Tag.FLUID)
    assert int(fluid.sum()) == 80000
    walls = ~fluid
    r_walls = state["r"][walls].clone()
    mass0 = float(state["mass"].double().sum())
    nb = fn.allocate(state["r"])
    assert adv._sph_fusable(nb) is not None
    dt = cfg.solver.dt
    assert dt < 1e-4
    st = state
    reallocs = 0
    front = [float(st["r"][fluid, 0].max())]
    for k in range(6000):
        st_, nb_ = adv(dt, st, nb)
        if bool(nb_.did_buffer_overflow):
            nb = fn.allocate(st["r"])
            reallocs += 1
            st, nb = adv(dt, st, nb)
            assert not bool(nb.did_buffer_overflow)
        else:
            st, nb = st_, nb_
        if (k + 1) % 100 == 0:
            front.append(float(st["r"][fluid, 0].max()))
            assert torch.equal(st["r"][walls], r_walls), k
            for f in ("r", "u", "rho", "p"):
                assert torch.isfinite(st[f]).all(), (k, f)
    front = np.array(front)
    print(f"dam break dx=0.005: {reallocs} re-allocations; front {front[0]:.4f} -> "
          f"{front[-1]:.4f}; largest front retreat {float(np.max(front[:-1] - front[1:])):.2e}",
          flush=True)
    assert float(st["mass"].double().sum()) == mass0
    assert (np.diff(front) >= -1e-6).all(), front
    assert front[-1] - front[0] > 0.1
    rho = st["rho"][fluid]
    assert 0.9 < float(rho.min()) and float(rho.max()) < 1.2


def test_dam_break_oracle_parity_once_the_column_moves():
    """dx = 0.04 with the case file's dt = 3e-4: 1000 steps on the GPU (t =
    0.3, the front has left the column), then 50 more from that state on the
    GPU and in the float32/float64 oracles."""
    cfg, box, state, fn, adv, solver, bc = _dam_break(0.04, 0.0003)
    nb = fn.allocate(state["r"])
    st = state
    fluid = state["tag"] == int(Tag.FLUID)
    x0 = float(state["r"][fluid, 0].max())
    for _ in range(1000):
        st, nb = adv(cfg.solver.dt, st, nb)
    assert not bool(nb.did_buffer_overflow)
    moved = float(st["r"][fluid, 0].max()) - x0
    print(f"dam break dx=0.04: front moved {moved:.4f} in 1000 steps", flush=True)
    assert moved > 0.02  # the column is moving
    start = {k: v.clone() for k, v in st.items()}
    for _ in range(50):
        st, nb = adv(cfg.solver.dt, st, nb)
    assert not bool(nb.did_buffer_overflow)
    R = G.OracleRef(cfg, box, solver, bc)
    ref, ref64 = R.run(start, 50), R.run(start, 50, f64=True)
    G.assert_within_noise_floor(st, ref, ref64, 1e-5, fields=("r", "u", "v", "rho"),
                                label="db 1000+50")
    G.assert_within_noise_floor(st, ref, ref64, 1e-5, fields=("dudt", "p"), factor=8.0,
                                label="db 1000+50")


def test_ensemble_64_members_largest_equal_their_own_steps():
    """The benchmark's 64-member ensemble (32 Taylor-Green vortices, 32
    lid-driven cavities, dx drawn from {0.025, 0.02, 0.0125, 0.01, 0.005}):
    20 batched steps, graph-replayed as the benchmark runs them; the four
    largest members equal their own fused steps bitwise."""
    import bench
    from jax_sph_amd.ensemble import Ensemble, setup_member
    cfgs, _ = bench.ensemble_configs(64, 0)
    ens = Ensemble.from_configs(cfgs)
    info = ens.info()
    assert info["members"] == 64 and info["particles"] == 872572
    ens.capture(10)
    ens.replay(2)
    torch.cuda.synchronize()
    assert not ens.did_buffer_overflow
    sizes = [ens.state(k)["r"].shape[0] for k in range(64)]
    largest = list(np.argsort(sizes, kind="stable")[::-1][:4])
    fresh, _ = bench.ensemble_configs(64, 0)
    for k in largest:
        adv, st, nb, dt = setup_member(fresh[k])
        for _ in range(20):
            st, nb = adv(dt, st, nb)
        torch.cuda.synchronize()
        got = ens.state(k)
        for f in ("r", "u", "v", "dudt", "dvdt", "rho", "p", "drhodt"):
            a, b = got[f].cpu().numpy(), st[f].cpu().numpy()
            assert np.array_equal(a, b), (k, sizes[k], f, np.abs(a - b).max())
        ia, ib = ens.error_info(k).cpu().numpy(), nb.info.cpu().numpy()
        assert np.array_equal(ia[[0, 1, 2, 3]], ib[[0, 1, 2, 3]]), (k, ia, ib)

"""Hand-written adjoint of the si_euler step (sph_step_vjp) and the
checkpointed differentiable rollout, against central finite differences of the
float64 oracle rollout (the reference computes these gradients with jax.grad,
which is not available here; finite differences of the reference algorithm
pin the same derivative)."""
import numpy as np
import pytest
import torch

from oracle import nl_oracle as nlo
from oracle import sph_oracle as so

pytestmark = pytest.mark.gpu

js = pytest.importorskip("jax_sph_amd")
from jax_sph_amd import adjoint, partition  # noqa: E402
from tests import gpu_helpers as G  # noqa: E402


def _setup(kernel="QSK", dx=0.05, seed=0):
    ov = {"case": {"dim": 2, "dx": dx}, "kernel": {"name": kernel}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("tgv", ov)
    gen = torch.Generator().manual_seed(seed)
    n = state["r"].shape[0]
    st = dict(state)
    st["u"] = (state["u"].cpu() + 0.05 * torch.randn(n, 2, generator=gen)).cuda()
    st["r"] = shift(state["r"], (0.1 * dx * torch.randn(n, 2, generator=gen)).cuda())
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    return cfg, box, st, solver, fn, shift, gen


def _oracle_rollout(st_np, solver, box, dt, n_steps):
    ns = nlo.NeighborSearch(box, solver._kernel_fn.cutoff, mask_self=False, dtype=np.float64)
    adv = so.si_euler(0.0, G.oracle_solver64(solver, box).forward, box, lambda s: s,
                      ns.update, np.float64)
    st = G.to_f64(st_np)
    nb = ns.allocate(st["r"])
    for _ in range(n_steps):
        st, nb = adv(dt, st, nb)
    return st


@pytest.mark.parametrize("kernel,n_steps,k", [
    ("QSK", 1, 1), ("QSK", 6, 2), ("QSK", 6, 6), ("WC2K", 5, 2),
])
def test_rollout_gradient_matches_oracle_fd(kernel, n_steps, k):
    cfg, box, st, solver, fn, shift, gen = _setup(kernel)
    dt = cfg.solver.dt
    n = st["r"].shape[0]
    w_u = torch.randn(n, 2, generator=gen).cuda()
    w_r = torch.randn(n, 2, generator=gen).cuda()
    u0 = st["u"].clone().requires_grad_(True)
    r0 = st["r"].clone().requires_grad_(True)
    st_in = dict(st)
    st_in["u"], st_in["r"] = u0, r0
    u, r, runner = adjoint.rollout(st_in, solver, fn, shift, dt, n_steps, checkpoint_every=k)
    loss = (w_u * u).sum() + (w_r * r).sum()
    loss.backward()
    gu, gr = u0.grad.double().cpu().numpy(), r0.grad.double().cpu().numpy()
    assert np.isfinite(gu).all() and np.isfinite(gr).all()
    base = G.to_np(st)
    wu, wr = w_u.double().cpu().numpy(), w_r.double().cpu().numpy()

    def L(eps, du, dr):
        s = dict(base)
        s = G.to_f64(s)
        s["u"] = s["u"] + eps * du
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, solver, box, dt, n_steps)
        return float((wu * out["u"]).sum() + (wr * out["r"]).sum())

    rng = np.random.default_rng(1)
    eps = 1e-6
    for trial in range(3):
        du = rng.standard_normal((n, 2))
        dr = rng.standard_normal((n, 2)) * (0.0 if trial == 0 else 1.0)
        fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
        ad = float((gu * du).sum() + (gr * dr).sum())
        assert abs(ad - fd) <= 1e-3 * abs(fd) + 1e-6, (trial, ad, fd)


def test_checkpoint_interval_does_not_change_the_gradient():
    cfg, box, st, solver, fn, shift, gen = _setup()
    n = st["r"].shape[0]
    w = torch.randn(n, 2, generator=gen).cuda()
    grads = []
    for k in (1, 3, 9):
        u0 = st["u"].clone().requires_grad_(True)
        s = dict(st)
        s["u"] = u0
        u, r, _ = adjoint.rollout(s, solver, fn, shift, cfg.solver.dt, 9, checkpoint_every=k)
        (w * u).sum().backward()
        grads.append(u0.grad.clone())
    assert torch.equal(grads[0], grads[1]) and torch.equal(grads[0], grads[2])


def test_unsupported_configurations_raise():
    cfg, box, st, solver, fn, shift, gen = _setup()
    with pytest.raises(NotImplementedError):
        adjoint.rollout(st, solver, fn, shift, cfg.solver.dt, 2, tvf=1.0)
    skin_fn = partition.neighbor_list(solver.displacement_fn, box,
                                      r_cutoff=solver._kernel_fn.cutoff, mask_self=False,
                                      format=partition.Sparse)
    with pytest.raises(ValueError):
        adjoint.rollout(st, solver, skin_fn, shift, cfg.solver.dt, 2)

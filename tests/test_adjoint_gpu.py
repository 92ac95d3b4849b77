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
from jax_sph_amd import adjoint, integrator, partition  # noqa: E402
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


def _oracle_rollout(st_np, solver, box, dt, n_steps, tvf=0.0, bc_fn=None):
    ns = nlo.NeighborSearch(box, solver._kernel_fn.cutoff, mask_self=False, dtype=np.float64)
    bcf = G.oracle_bc(bc_fn, solver.dim) if bc_fn is not None else (lambda s: s)
    adv = so.si_euler(tvf, G.oracle_solver64(solver, box).forward, box, bcf, ns.update,
                      np.float64)
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


def _setup_tvf(dim, kernel, n_side, seed=0):
    """Taylor-Green vortex with the transport-velocity formulation (tvf = 1):
    2D on the unit square or 3D on the (2 pi)^3 cube, with jittered
    positions, a perturbed velocity and a nonzero initial dvdt (so the first
    step's v = u + c dvdt differs from u)."""
    dx = (2 * np.pi if dim == 3 else 1.0) / n_side
    ov = {"case": {"dim": dim, "dx": dx}, "kernel": {"name": kernel}, "solver": {"tvf": 1.0}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("tgv", ov)
    gen = torch.Generator().manual_seed(seed)
    n = state["r"].shape[0]
    st = dict(state)
    st["u"] = (state["u"].cpu() + 0.05 * torch.randn(n, dim, generator=gen)).cuda()
    st["r"] = shift(state["r"], (0.1 * dx * torch.randn(n, dim, generator=gen)).cuda())
    st["dvdt"] = (10.0 * torch.randn(n, dim, generator=gen)).cuda()
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    return cfg, box, st, solver, fn, shift, gen


@pytest.mark.parametrize("dim,kernel,n_side,n_steps,k", [
    (2, "QSK", 20, 1, 1), (2, "QSK", 20, 5, 2), (2, "WC2K", 20, 4, 4),
    (3, "QSK", 10, 1, 1), (3, "QSK", 10, 4, 2), (3, "WC2K", 12, 3, 3),
])
def test_tvf_rollout_gradient_matches_oracle_fd(dim, kernel, n_side, n_steps, k):
    """The transport-velocity step (tvf = 1: the A = rho u (v - u) stress in
    dudt and the background-pressure dvdt, solver.py:199-256, carried through
    v = u + tvf dt/2 dvdt, integrator.py:26-30): directional derivatives of a
    linear functional of the final (u, r) with respect to the initial u and r
    against central differences of the float64 oracle rollout, in 2D and 3D."""
    cfg, box, st, solver, fn, shift, gen = _setup_tvf(dim, kernel, n_side)
    dt = cfg.solver.dt
    n = st["r"].shape[0]
    w_u = torch.randn(n, dim, generator=gen).cuda()
    w_r = torch.randn(n, dim, generator=gen).cuda()
    u0 = st["u"].clone().requires_grad_(True)
    r0 = st["r"].clone().requires_grad_(True)
    st_in = dict(st)
    st_in["u"], st_in["r"] = u0, r0
    u, r, runner = adjoint.rollout(st_in, solver, fn, shift, dt, n_steps, tvf=1.0,
                                   checkpoint_every=k)
    ((w_u * u).sum() + (w_r * r).sum()).backward()
    gu, gr = u0.grad.double().cpu().numpy(), r0.grad.double().cpu().numpy()
    assert np.isfinite(gu).all() and np.isfinite(gr).all()
    base = G.to_np(st)
    wu, wr = w_u.double().cpu().numpy(), w_r.double().cpu().numpy()

    def L(eps, du, dr, tvf=1.0):
        s = G.to_f64(dict(base))
        s["u"] = s["u"] + eps * du
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, solver, box, dt, n_steps, tvf=tvf)
        return float((wu * out["u"]).sum() + (wr * out["r"]).sum())

    rng = np.random.default_rng(1)
    eps = 1e-6
    for trial in range(3):
        du = rng.standard_normal((n, dim))
        dr = rng.standard_normal((n, dim)) * (0.0 if trial == 0 else 1.0)
        fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
        ad = float((gu * du).sum() + (gr * dr).sum())
        assert abs(ad - fd) <= 1e-3 * abs(fd) + 1e-6, (trial, ad, fd)
        if n_steps >= 2 and trial == 2:
            # the transport-velocity terms move this derivative by far more
            # than the tolerance: a transpose without them would fail
            fd0 = (L(eps, du, dr, 0.0) - L(-eps, du, dr, 0.0)) / (2 * eps)
            assert abs(fd - fd0) > 0.02 * abs(fd), (fd, fd0)


def test_tvf_graph_rollout_equals_eager_rollout():
    """GraphRollout of the transport-velocity step (its checkpoints carry
    dvdt) gives the eager rollout's values and gradients bit for bit."""
    cfg, box, st, solver, fn, shift, gen = _setup_tvf(3, "QSK", 10)
    n = st["r"].shape[0]
    w = torch.randn(n, 3, generator=gen).cuda()
    u0 = st["u"].clone().requires_grad_(True)
    s = dict(st)
    s["u"] = u0
    u, r, _ = adjoint.rollout(s, solver, fn, shift, cfg.solver.dt, 6, tvf=1.0,
                              checkpoint_every=3)
    ((w * u).sum() + (w * r).sum()).backward()
    gr_ = adjoint.GraphRollout(st, solver, fn, shift, cfg.solver.dt, 6, tvf=1.0,
                               checkpoint_every=3)
    for _ in range(2):
        uu = st["u"].clone().requires_grad_(True)
        ug, rg = gr_(uu, st["r"])
        ((w * ug).sum() + (w * rg).sum()).backward()
        torch.cuda.synchronize()
        assert torch.equal(ug, u) and torch.equal(rg, r)
        assert torch.equal(uu.grad, u0.grad)


@pytest.mark.parametrize("case", ["tgv", "ldc"])
def test_reference_gradient_check_configuration(case):
    """The reference's own gradient validation (notebooks/iclr24_grads.ipynb,
    cell 5): dx = 1/20, viscosity 0.1, 10 warm-up steps, then the gradient of
    0.5 sum(u^2) after 5 more steps with respect to the positions, against
    finite differences.  Its WCSPH call passes is_bc_trick in the eta_limiter
    slot, so the cavity runs without the wall boundary treatment: its walls
    are plain particles held by the case hook (walls at rest, lid at u = 1).
    Here the differences come from the float64 oracle (central, eps = 1e-6)
    along random directions, to 1e-3 relative."""
    ov = {"case": {"dx": 0.05, "viscosity": 0.1}, "solver": {"t_end": 1.0}}
    if case == "tgv":
        ov["case"]["dim"] = 2
    else:
        ov["solver"]["is_bc_trick"] = False
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup(case, ov)
    assert not solver.is_bc_trick
    if case == "ldc":
        assert bc.kind == 1 and (state["tag"] != 0).any()
    dt = cfg.solver.dt
    cut = solver._kernel_fn.cutoff
    warm = partition.neighbor_list(disp, box, r_cutoff=cut, mask_self=False,
                                   format=partition.Sparse)
    nb = warm.allocate(state["r"])
    adv = integrator.si_euler(0.0, solver.forward_wrapper(), shift, bc, None)
    st = state
    for _ in range(10):
        st, nb = adv(dt, st, nb)
    fn = partition.neighbor_list(disp, box, r_cutoff=cut, capacity_multiplier=1.25,
                                 mask_self=False, format=partition.Sparse, skin=0.0)
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["r"] = r0
    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, 5, checkpoint_every=5, bc_fn=bc)
    (0.5 * (u.double() ** 2).sum()).backward()
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gr).all()
    base = G.to_np(st)
    n, dim = base["r"].shape

    def L(eps, dr):
        s = G.to_f64(dict(base))
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, solver, box, dt, 5, bc_fn=bc)
        return 0.5 * float((out["u"] ** 2).sum())

    # along the gradient itself (no cancellation in the directional sum), and
    # along random directions, whose sum over 2N terms cancels to about
    # 1/sqrt(2N) of sum |g_i d_i|: there the float32 adjoint is held to that scale
    rng = np.random.default_rng(2)
    eps = 1e-6
    dirs = [gr / np.linalg.norm(gr)] + [rng.standard_normal((n, dim)) for _ in range(3)]
    for trial, dr in enumerate(dirs):
        fd = (L(eps, dr) - L(-eps, dr)) / (2 * eps)
        ad = float((gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gr * dr).sum())
        assert abs(ad - fd) <= 1e-3 * scale, (case, trial, ad, fd, scale)


@pytest.mark.parametrize("tvf,n_steps,k", [(0.0, 1, 1), (0.0, 4, 2), (1.0, 3, 3)])
def test_wall_bc_rollout_gradient_matches_oracle_fd(tvf, n_steps, k):
    """The lid-driven cavity as its case file runs it: the generalized wall
    boundary condition (is_bc_trick, solver.py:431-530, no-slip walls) sets
    the walls' velocity, pressure and density from Shepard averages of their
    fluid neighbours before every force evaluation, and the case hook holds
    the walls.  dx = 0.05, viscosity 0.1, 10 warm-up steps, then the gradient
    of 0.5 sum(u^2) + <w, r> after n steps in the fluid's initial velocity and
    in all positions, against float64 oracle central differences (eps = 1e-6):
    along the gradient to 1e-3 relative, along random directions to 1e-3 of
    sum |g_i d_i|.  The same differences taken without the wall boundary
    condition differ by more than 1%, ten times the tolerance (from two steps
    on: after one step u' and r' depend on the forces of the state before), so
    the wall terms of the transpose are exercised.  (Walls' initial velocities are the hook's
    constants: their cotangent is zero, and they are not perturbed.)"""
    ov = {"case": {"dx": 0.05, "viscosity": 0.1}, "solver": {"t_end": 1.0, "tvf": tvf}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("ldc", ov)
    assert solver.is_bc_trick and not solver.is_free_slip and bc.kind == 1
    dt = cfg.solver.dt
    cut = solver._kernel_fn.cutoff
    warm = partition.neighbor_list(disp, box, r_cutoff=cut, mask_self=False,
                                   format=partition.Sparse)
    nb = warm.allocate(state["r"])
    adv = integrator.si_euler(tvf, solver.forward_wrapper(), shift, bc, None)
    st = state
    for _ in range(10):
        st, nb = adv(dt, st, nb)
    fn = partition.neighbor_list(disp, box, r_cutoff=cut, capacity_multiplier=1.25,
                                 mask_self=False, format=partition.Sparse, skin=0.0)
    n = st["r"].shape[0]
    gen = torch.Generator().manual_seed(5)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    u0 = st["u"].clone().requires_grad_(True)
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["u"], s_in["r"] = u0, r0
    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, n_steps, tvf=tvf,
                              checkpoint_every=k, bc_fn=bc)
    (0.5 * (u.double() ** 2).sum() + (w * r.double()).sum()).backward()
    # the graph-replayed rollout (its transposes read the records its forward
    # left, after the wall boundary condition) gives the same bits
    gro = adjoint.GraphRollout(st, solver, fn, shift, dt, n_steps, tvf=tvf,
                               checkpoint_every=k, bc_fn=bc)
    u2 = st["u"].clone().requires_grad_(True)
    r2 = st["r"].clone().requires_grad_(True)
    ug, rg = gro(u2, r2)
    (0.5 * (ug.double() ** 2).sum() + (w * rg.double()).sum()).backward()
    assert torch.equal(ug, u) and torch.equal(rg, r)
    assert torch.equal(u2.grad, u0.grad) and torch.equal(r2.grad, r0.grad)
    fluid = (st["tag"] == 0).cpu().numpy()[:, None]
    gu = u0.grad.double().cpu().numpy() * fluid
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gu).all() and np.isfinite(gr).all()
    base = G.to_np(st)
    wn = w.cpu().numpy()
    nobc = G.setup("ldc", {**ov, "solver": {**ov["solver"], "is_bc_trick": False}})[-1]

    def L(eps, du, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["u"] = s["u"] + eps * du
        s["v"] = s["u"].copy()   # the rollout starts from v = u (v' is recomputed)
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, sv, box, dt, n_steps, tvf=tvf, bc_fn=bc)
        return 0.5 * float((out["u"] ** 2).sum()) + float((wn * out["r"]).sum())

    rng = np.random.default_rng(3)
    eps = 1e-6
    norm = np.sqrt((gu ** 2).sum() + (gr ** 2).sum())
    dirs = [(gu / norm, gr / norm)] + [(rng.standard_normal(gu.shape) * fluid,
                                        rng.standard_normal(gr.shape)) for _ in range(2)]
    for trial, (du, dr) in enumerate(dirs):
        fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
        ad = float((gu * du).sum() + (gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gu * du).sum() + np.abs(gr * dr).sum())
        assert abs(ad - fd) <= 1e-3 * scale, (trial, ad, fd, scale)
        if trial == 0 and n_steps >= 2:   # one step: u', r' do not see the new forces
            fd_nobc = (L(eps, du, dr, nobc) - L(-eps, du, dr, nobc)) / (2 * eps)
            assert abs(fd_nobc - fd) > 0.01 * abs(fd), (fd, fd_nobc)


@pytest.mark.parametrize("case,tvf,n_steps,k", [("tgv", 0.0, 1, 1), ("tgv", 0.0, 4, 2),
                                                 ("tgv", 1.0, 3, 3), ("ldc", 0.0, 12, 4),
                                                 ("ldc_gwbc", 0.0, 12, 4),
                                                 ("ldc_gwbc", 1.0, 12, 3)])
def test_continuity_rollout_gradient_matches_oracle_fd(case, tvf, n_steps, k):
    """Density continuity (density_evolution, solver.py:24-30): the density
    is advanced from the one the step is given, so it is carried, and its
    cotangent crosses the steps.  2D TGV (dx = 0.05) and the lid-driven cavity
    with plain walls held by the hook (dx = 0.05, 10 warm-up steps); the
    gradient of 0.5 sum(u^2) + <w, r> after n steps in the initial velocity
    (fluid) and positions, against float64 oracle central differences, to
    1e-3 as above.  From two steps on, the differences of the same rollouts
    with density summation differ by more than 1% (the cavity, whose density
    moves slowly, over 12 steps).  The cavity also runs with the wall
    boundary condition (ldc_gwbc): the walls' rho^ is then carried to the
    next step's continuity update of their fluid neighbours."""
    ov = {"case": {"dx": 0.05, "viscosity": 0.1}, "solver": {"t_end": 1.0, "tvf": tvf,
                                                             "density_evolution": True}}
    gwbc = case == "ldc_gwbc"
    case = "ldc" if gwbc else case
    if case == "tgv":
        ov["case"]["dim"] = 2
    else:
        ov["solver"]["is_bc_trick"] = gwbc
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup(case, ov)
    assert solver.is_rho_evol and solver.is_bc_trick == gwbc
    dt = cfg.solver.dt
    cut = solver._kernel_fn.cutoff
    st = state
    if case == "ldc":
        warm = partition.neighbor_list(disp, box, r_cutoff=cut, mask_self=False,
                                       format=partition.Sparse)
        nb = warm.allocate(state["r"])
        adv = integrator.si_euler(tvf, solver.forward_wrapper(), shift, bc, None)
        for _ in range(10):
            st, nb = adv(dt, st, nb)
    fn = partition.neighbor_list(disp, box, r_cutoff=cut, capacity_multiplier=1.25,
                                 mask_self=False, format=partition.Sparse, skin=0.0)
    n = st["r"].shape[0]
    gen = torch.Generator().manual_seed(7)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    u0 = st["u"].clone().requires_grad_(True)
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["u"], s_in["r"] = u0, r0
    bc_fn = bc if case == "ldc" else None
    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, n_steps, tvf=tvf,
                              checkpoint_every=k, bc_fn=bc_fn)
    (0.5 * (u.double() ** 2).sum() + (w * r.double()).sum()).backward()
    gro = adjoint.GraphRollout(st, solver, fn, shift, dt, n_steps, tvf=tvf,
                               checkpoint_every=k, bc_fn=bc_fn)
    u2 = st["u"].clone().requires_grad_(True)
    r2 = st["r"].clone().requires_grad_(True)
    ug, rg = gro(u2, r2)
    (0.5 * (ug.double() ** 2).sum() + (w * rg.double()).sum()).backward()
    assert torch.equal(ug, u) and torch.equal(rg, r)
    assert torch.equal(u2.grad, u0.grad) and torch.equal(r2.grad, r0.grad)
    fluid = (st["tag"] == 0).cpu().numpy()[:, None]
    gu = u0.grad.double().cpu().numpy() * fluid
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gu).all() and np.isfinite(gr).all()
    base = G.to_np(st)
    wn = w.cpu().numpy()
    summ = G.setup(case, {**ov, "solver": {**ov["solver"], "density_evolution": False}})[-1]

    def L(eps, du, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["u"] = s["u"] + eps * du
        s["v"] = s["u"].copy()
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, sv, box, dt, n_steps, tvf=tvf, bc_fn=bc_fn)
        return 0.5 * float((out["u"] ** 2).sum()) + float((wn * out["r"]).sum())

    rng = np.random.default_rng(4)
    eps = 1e-6
    norm = np.sqrt((gu ** 2).sum() + (gr ** 2).sum())
    dirs = [(gu / norm, gr / norm)] + [(rng.standard_normal(gu.shape) * fluid,
                                        rng.standard_normal(gr.shape)) for _ in range(2)]
    for trial, (du, dr) in enumerate(dirs):
        fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
        ad = float((gu * du).sum() + (gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gu * du).sum() + np.abs(gr * dr).sum())
        assert abs(ad - fd) <= 1e-3 * scale, (trial, ad, fd, scale)
        if trial == 0 and n_steps >= 2:
            fd_s = (L(eps, du, dr, summ) - L(-eps, du, dr, summ)) / (2 * eps)
            assert abs(fd_s - fd) > 0.01 * abs(fd), (fd, fd_s)


class _SmoothForce:
    """g(r) = a (sin 2 pi y, cos 2 pi x): a body force whose derivative in r
    the adjoint must carry; ``np`` is its float64 twin for the oracle."""

    def __init__(self, a):
        self.a = float(a)

    def __call__(self, r):
        t = 2 * np.pi
        return self.a * torch.stack([torch.sin(t * r[:, 1]), torch.cos(t * r[:, 0])], 1)

    def np(self, r):
        t = 2 * np.pi
        return self.a * np.stack([np.sin(t * r[:, 1]), np.cos(t * r[:, 0])], 1)


@pytest.mark.parametrize("case,n_steps,k", [("pf", 3, 2), ("smooth", 1, 1), ("smooth", 4, 2),
                                            ("smooth_cont", 4, 2)])
def test_body_force_rollout_gradient_matches_oracle_fd(case, n_steps, k, monkeypatch):
    """A position-dependent body force g(r) (g_ext_fn, solver.py:936).  The
    Poiseuille channel as its case file runs it (pf.py:134-143: the force
    switched on between the walls by comparisons, so dg/dr = 0 as under
    jax.grad; no-slip walls with the wall boundary condition, held by the
    hook; dx = 0.05, fluid velocities perturbed), and a 2D TGV (dx = 0.05)
    driven by the smooth force a (sin 2 pi y, cos 2 pi x), a = 200, whose
    dg/dr the rollout takes by autograd through the force (smooth_cont: with
    density continuity, whose transpose reads the records the step left).  Gradient of
    0.5 sum(u^2) + <w, r> after n steps in the initial velocity (fluid) and
    positions against float64 oracle central differences, to 1e-3 as above.
    For the smooth force, dropping the dg/dr term moves the directional
    derivative by more than the tolerance (from two steps on)."""
    if case == "pf":
        ov = {"case": {"dx": 0.05}}
        cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("pf", ov)
        assert solver.g_const is None and solver.is_bc_trick and bc.kind == 1
        bc_fn = bc
    else:
        ov = {"case": {"dim": 2, "dx": 0.05, "viscosity": 0.1},
              "solver": {"t_end": 1.0, "density_evolution": case == "smooth_cont"}}
        cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("tgv", ov)
        from jax_sph_amd import case_setup as cs
        solver = cs.make_solver(cfg, disp, eos, _SmoothForce(200.0))
        bc_fn = None
    dt = cfg.solver.dt
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(11)
    fluid_t = (state["tag"] == 0)[:, None]
    st = dict(state)
    st["u"] = torch.where(fluid_t, state["u"] + 0.05 * torch.randn(n, 2, generator=gen).cuda(),
                          state["u"])
    st["v"] = st["u"].clone()
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    w = torch.randn(n, 2, generator=gen).cuda().double()

    def grads():
        u0 = st["u"].clone().requires_grad_(True)
        r0 = st["r"].clone().requires_grad_(True)
        s_in = dict(st)
        s_in["u"], s_in["r"] = u0, r0
        u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, n_steps, checkpoint_every=k,
                                  bc_fn=bc_fn)
        (0.5 * (u.double() ** 2).sum() + (w * r.double()).sum()).backward()
        return u0.grad.double().cpu().numpy(), r0.grad.double().cpu().numpy()

    fluid = fluid_t.cpu().numpy()
    gu, gr = grads()
    gu = gu * fluid
    assert np.isfinite(gu).all() and np.isfinite(gr).all()
    if case != "pf":
        with monkeypatch.context() as m:
            m.setattr(adjoint._Runner, "_body_force_pullback", lambda self, s_, c_: None)
            gu_no, gr_no = grads()
        gu_no = gu_no * fluid
    base = G.to_np(st)
    wn = w.cpu().numpy()

    def L(eps, du, dr):
        s = G.to_f64(dict(base))
        s["u"] = s["u"] + eps * du
        s["v"] = s["u"].copy()
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, solver, box, dt, n_steps, bc_fn=bc_fn)
        return 0.5 * float((out["u"] ** 2).sum()) + float((wn * out["r"]).sum())

    rng = np.random.default_rng(6)
    eps = 1e-6
    norm = np.sqrt((gu ** 2).sum() + (gr ** 2).sum())
    dirs = [(gu / norm, gr / norm)] + [(rng.standard_normal(gu.shape) * fluid,
                                        rng.standard_normal(gr.shape)) for _ in range(2)]
    for trial, (du, dr) in enumerate(dirs):
        fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
        ad = float((gu * du).sum() + (gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gu * du).sum() + np.abs(gr * dr).sum())
        msg = (case, trial, ad, fd, scale)
        if case != "pf":
            ad_no = float((gu_no * du).sum() + (gr_no * dr).sum())
            msg += (ad_no,)
            print(f"body force {case} n={n_steps} trial {trial}: ad {ad:.9g} fd {fd:.9g} "
                  f"without dg/dr {ad_no:.9g} scale {scale:.6g}", flush=True)
        assert abs(ad - fd) <= 1e-3 * scale, msg
        if case != "pf" and n_steps >= 2 and trial == 0:
            assert abs(ad_no - fd) > 2e-3 * scale, msg


@pytest.mark.parametrize("case,tvf,n_steps,k", [("tgv", 0.0, 1, 1), ("tgv", 0.0, 4, 2),
                                                 ("tgv", 1.0, 3, 3), ("db", 0.0, 2, 1),
                                                 ("db", 0.0, 4, 2)])
def test_artificial_viscosity_rollout_gradient_matches_oracle_fd(case, tvf, n_steps, k):
    """Artificial viscosity (solver.py:404-428) on fluid-fluid pairs.  The 2D
    TGV (dx = 0.05, density summation, alpha = 0.5) and the dam break as its
    case file runs it (cases/db.yaml: density continuity, the wall boundary
    condition with gravity normal to the bottom wall, alpha = 0.1, the hook
    holding the walls; dx = 0.1, 50 warm-up steps so the column moves).  TGV:
    the gradient of 0.5 sum(u^2) + <w, r> after n steps in the fluid's
    initial velocity and all positions; dam break: the gradient of <w, u>
    after n steps in all initial positions, along position directions only
    (so the derivative is the dynamics' alone: along velocity directions the
    identity part of <w, u> hides it; after one step u' does not depend on
    the positions).  Against float64 oracle central
    differences, to 1e-3 as above.  From two steps on, the TGV differences
    taken without the artificial viscosity differ by more than 1%, so its
    transpose is exercised (in the dam break, whose velocities are still
    nearly uniform after the warm-up, it moves them too little to test; there
    the differences without the wall boundary condition differ by more than
    1%).  GraphRollout equals the eager rollout bit for bit."""
    if case == "tgv":
        ov = {"case": {"dim": 2, "dx": 0.05, "viscosity": 0.01},
              "solver": {"t_end": 1.0, "tvf": tvf, "artificial_alpha": 0.5}}
    else:
        ov = {"case": {"dx": 0.1}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup(case, ov)
    assert solver.artificial_alpha > 0
    if case == "db":
        assert solver.is_rho_evol and solver.is_bc_trick and bc.kind == 2
    dt = cfg.solver.dt
    cut = solver._kernel_fn.cutoff
    st = state
    bc_fn = bc if case == "db" else None
    if case == "db":
        warm = partition.neighbor_list(disp, box, r_cutoff=cut, mask_self=False,
                                       format=partition.Sparse)
        nb = warm.allocate(state["r"])
        adv = integrator.si_euler(tvf, solver.forward_wrapper(), shift, bc, None)
        for _ in range(50):
            st, nb = adv(dt, st, nb)
    else:
        gen0 = torch.Generator().manual_seed(2)
        n0 = st["r"].shape[0]
        st = dict(st)
        st["u"] = (st["u"].cpu() + 0.1 * torch.randn(n0, 2, generator=gen0)).cuda()
        st["v"] = st["u"].clone()
    fn = partition.neighbor_list(disp, box, r_cutoff=cut, capacity_multiplier=1.25,
                                 mask_self=False, format=partition.Sparse, skin=0.0)
    n = st["r"].shape[0]
    gen = torch.Generator().manual_seed(9)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    u0 = st["u"].clone().requires_grad_(True)
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["u"], s_in["r"] = u0, r0
    lin = case == "db"

    def loss(u, r):
        if lin:
            return (w * u.double()).sum() + 0.0 * r.sum()
        return 0.5 * (u.double() ** 2).sum() + (w * r.double()).sum()

    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, n_steps, tvf=tvf,
                              checkpoint_every=k, bc_fn=bc_fn)
    loss(u, r).backward()
    gro = adjoint.GraphRollout(st, solver, fn, shift, dt, n_steps, tvf=tvf,
                               checkpoint_every=k, bc_fn=bc_fn)
    u2 = st["u"].clone().requires_grad_(True)
    r2 = st["r"].clone().requires_grad_(True)
    ug, rg = gro(u2, r2)
    loss(ug, rg).backward()
    assert torch.equal(ug, u) and torch.equal(rg, r)
    assert torch.equal(u2.grad, u0.grad) and torch.equal(r2.grad, r0.grad)
    fluid = (st["tag"] == 0).cpu().numpy()[:, None]
    gu = u0.grad.double().cpu().numpy() * fluid
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gu).all() and np.isfinite(gr).all()
    base = G.to_np(st)
    wn = w.cpu().numpy()
    off = {"artificial_alpha": 0.0} if case == "tgv" else {"is_bc_trick": False}
    alt = G.setup(case, {**ov, "solver": {**ov.get("solver", {}), **off}})[-1]

    def L(eps, du, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["u"] = s["u"] + eps * du
        s["v"] = s["u"].copy()
        s["r"] = np.mod(s["r"] + eps * dr, box)
        s["r"] = np.where(s["r"] >= box, s["r"] - box, s["r"])  # -tiny mod box rounds to box
        out = _oracle_rollout(s, sv, box, dt, n_steps, tvf=tvf, bc_fn=bc_fn)
        if lin:
            return float((wn * out["u"]).sum())
        return 0.5 * float((out["u"] ** 2).sum()) + float((wn * out["r"]).sum())

    rng = np.random.default_rng(8)
    eps = 1e-6
    if lin:
        gu = np.zeros_like(gu)
    norm = np.sqrt((gu ** 2).sum() + (gr ** 2).sum())
    dirs = [(gu / norm, gr / norm)] + [(rng.standard_normal(gu.shape) * fluid * (not lin),
                                        rng.standard_normal(gr.shape)) for _ in range(2)]
    for trial, (du, dr) in enumerate(dirs):
        fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
        ad = float((gu * du).sum() + (gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gu * du).sum() + np.abs(gr * dr).sum())
        print(f"av {case} tvf={tvf} n={n_steps} trial {trial}: ad {ad:.9g} fd {fd:.9g} "
              f"scale {scale:.6g}", flush=True)
        assert abs(ad - fd) <= 1e-3 * scale, (case, trial, ad, fd, scale)
        if trial == 0 and n_steps >= 2:
            fd_no = (L(eps, du, dr, alt) - L(-eps, du, dr, alt)) / (2 * eps)
            print(f"  without {off}: fd {fd_no:.9g}", flush=True)
            assert abs(fd_no - fd) > 0.01 * abs(fd), (fd, fd_no)


@pytest.mark.parametrize("case,tvf,n_steps,k", [("tgv", 0.0, 1, 1), ("tgv", 0.0, 4, 2),
                                                 ("tgv", 1.0, 3, 3), ("ldc", 0.0, 5, 1)])
def test_delta_sph_rollout_gradient_matches_oracle_fd(case, tvf, n_steps, k):
    """Delta-SPH with density summation (solver.py:259-311: standard SPH plus
    the velocity diffusion alpha h c_ref rho_ref V_j pi_ij grad W / rho_i over
    fluid neighbours; alpha = 0.1): the 2D TGV (dx = 0.05) with and without
    the transport velocity, and the lid-driven cavity (dx = 0.1 at the CFL
    time step, its wall boundary condition, the hook holding the walls and
    the lid).  The
    gradient of 0.5 sum(u^2) + <w, r> after n steps in the fluid's initial
    velocity (perturbed) and all positions, against float64 oracle central
    differences to 1e-3 along the gradient and two random directions.  From
    two steps on, the differences taken with the standard solver (no
    velocity diffusion) differ by more than 1%, so the term's transpose is
    exercised.  GraphRollout equals the eager rollout bit for bit."""
    ov = {"case": {"dim": 2, "dx": 0.05 if case == "tgv" else 0.1},
          "solver": {"name": "DELTA", "diff_alpha": 0.1, "tvf": tvf}}
    if case == "tgv":
        ov["case"]["viscosity"] = 0.01
        ov["solver"]["t_end"] = 1.0
    else:
        ov["solver"]["dt"] = None
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup(case, ov)
    assert solver.solver == "DELTA" and not solver.is_rho_evol
    dt = cfg.solver.dt
    cut = solver._kernel_fn.cutoff
    bc_fn = bc if case == "ldc" else None
    gen0 = torch.Generator().manual_seed(2)
    n = state["r"].shape[0]
    fluid_t = (state["tag"] == 0).cpu()[:, None]
    st = dict(state)
    st["u"] = (st["u"].cpu() + 0.1 * torch.randn(n, 2, generator=gen0) * fluid_t).cuda()
    st["v"] = st["u"].clone()
    fn = partition.neighbor_list(disp, box, r_cutoff=cut, capacity_multiplier=1.25,
                                 mask_self=False, format=partition.Sparse, skin=0.0)
    gen = torch.Generator().manual_seed(9)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    u0 = st["u"].clone().requires_grad_(True)
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["u"], s_in["r"] = u0, r0

    def loss(u, r):
        return 0.5 * (u.double() ** 2).sum() + (w * r.double()).sum()

    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, n_steps, tvf=tvf,
                              checkpoint_every=k, bc_fn=bc_fn)
    loss(u, r).backward()
    gro = adjoint.GraphRollout(st, solver, fn, shift, dt, n_steps, tvf=tvf,
                               checkpoint_every=k, bc_fn=bc_fn)
    u2 = st["u"].clone().requires_grad_(True)
    r2 = st["r"].clone().requires_grad_(True)
    ug, rg = gro(u2, r2)
    loss(ug, rg).backward()
    assert torch.equal(ug, u) and torch.equal(rg, r)
    assert torch.equal(u2.grad, u0.grad) and torch.equal(r2.grad, r0.grad)
    fluid = fluid_t.numpy()
    gu = u0.grad.double().cpu().numpy() * fluid
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gu).all() and np.isfinite(gr).all()
    base = G.to_np(st)
    wn = w.cpu().numpy()
    alt = G.setup(case, {**ov, "solver": {**ov["solver"], "name": "SPH"}})[-1]

    def L(eps, du, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["u"] = s["u"] + eps * du
        s["v"] = s["u"].copy()
        s["r"] = np.mod(s["r"] + eps * dr, box)
        s["r"] = np.where(s["r"] >= box, s["r"] - box, s["r"])  # -tiny mod box rounds to box
        out = _oracle_rollout(s, sv, box, dt, n_steps, tvf=tvf, bc_fn=bc_fn)
        return 0.5 * float((out["u"] ** 2).sum()) + float((wn * out["r"]).sum())

    rng = np.random.default_rng(8)
    eps = 1e-6
    norm = np.sqrt((gu ** 2).sum() + (gr ** 2).sum())
    dirs = [(gu / norm, gr / norm)] + [(rng.standard_normal(gu.shape) * fluid,
                                        rng.standard_normal(gr.shape)) for _ in range(2)]
    for trial, (du, dr) in enumerate(dirs):
        fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
        ad = float((gu * du).sum() + (gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gu * du).sum() + np.abs(gr * dr).sum())
        print(f"delta {case} tvf={tvf} n={n_steps} trial {trial}: ad {ad:.9g} fd {fd:.9g} "
              f"scale {scale:.6g}", flush=True)
        assert abs(ad - fd) <= 1e-3 * scale, (case, trial, ad, fd, scale)
        if trial == 0 and n_steps >= 2:
            fd_no = (L(eps, du, dr, alt) - L(-eps, du, dr, alt)) / (2 * eps)
            print(f"  with the standard solver: fd {fd_no:.9g}", flush=True)
            assert abs(fd_no - fd) > 0.01 * abs(fd), (fd, fd_no)


@pytest.mark.parametrize("free,n_steps,k", [(False, 2, 1), (True, 2, 1), (True, 4, 2)])
def test_channel_walls_gradient_matches_oracle_fd(free, n_steps, k):
    """The Poiseuille channel (dx = 0.05, walls at rest, fluid velocities
    perturbed) with no-slip and with free-slip walls (solver.py:505-510: the
    walls take only the normal part of the Shepard average of their fluid
    neighbours' velocity, so the cotangents they pass back are projected on
    the normal).  Its body force is switched on between the walls, so the
    walls' pressure extrapolation uses the walls' own g (zero), not their
    fluid neighbours'.  The gradient of <w, u> after n steps in the initial
    positions along position directions (the dynamics alone), against float64
    oracle central differences to 1e-5 (two and four steps leave the float32
    rollout's rounding far below that: measured 2e-8 to 1e-7); with
    free-slip walls the differences with no-slip walls differ by more than
    1e-4 (measured 0.25%), ten times that bound.  The cavity's free-slip
    lid, whose normals follow it, raises."""
    ov = {"case": {"dx": 0.05}, "solver": {"free_slip": free}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("pf", ov)
    assert solver.is_bc_trick and solver.is_free_slip == free and solver.g_const is None
    if free:
        assert float(state["nw"].abs().sum()) > 0
    noslip = G.setup("pf", {"case": {"dx": 0.05}})[-1]
    dt = cfg.solver.dt
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(12)
    fluid_t = (state["tag"] == 0)[:, None]
    st = dict(state)
    st["u"] = torch.where(fluid_t, state["u"] + 0.2 * torch.randn(n, 2, generator=gen).cuda(),
                          state["u"])
    st["v"] = st["u"].clone()
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["r"] = r0
    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, n_steps, checkpoint_every=k,
                              bc_fn=bc)
    ((w * u.double()).sum() + 0.0 * r.sum()).backward()
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gr).all() and np.abs(gr).max() > 0
    base = G.to_np(st)
    wn = w.cpu().numpy()

    def L(eps, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, sv, box, dt, n_steps, bc_fn=bc)
        return float((wn * out["u"]).sum())

    rng = np.random.default_rng(13)
    eps = 1e-6
    dirs = [gr / np.sqrt((gr ** 2).sum())] + [rng.standard_normal(gr.shape) for _ in range(2)]
    for trial, dr in enumerate(dirs):
        fd = (L(eps, dr) - L(-eps, dr)) / (2 * eps)
        ad = float((gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gr * dr).sum())
        print(f"free slip {free} n={n_steps} trial {trial}: ad {ad:.9g} fd {fd:.9g} "
              f"scale {scale:.6g}", flush=True)
        assert abs(ad - fd) <= 1e-5 * scale, (trial, ad, fd, scale)
        if trial == 0 and free:
            fd_ns = (L(eps, dr, noslip) - L(-eps, dr, noslip)) / (2 * eps)
            print(f"  no-slip walls: fd {fd_ns:.9g}", flush=True)
            assert abs(fd_ns - fd) > 1e-4 * abs(fd), (fd, fd_ns)
    lov = {"case": {"dx": 0.05}, "solver": {"free_slip": True}}
    lcfg, lbox, lstate, _, lbc, _, ldisp, lshift, lsolver = G.setup("ldc", lov)
    lfn = partition.neighbor_list(ldisp, lbox, r_cutoff=lsolver._kernel_fn.cutoff,
                                  mask_self=False, format=partition.Sparse, skin=0.0)
    with pytest.raises(NotImplementedError, match="moving walls"):
        adjoint.rollout(lstate, lsolver, lfn, lshift, lcfg.solver.dt, 2, bc_fn=lbc)


@pytest.mark.parametrize("tvf", [0.0, 1.0])
def test_wall_bc_3d_gradient_matches_oracle_fd(tvf):
    """The 3D lid-driven cavity (dx = 0.05, periodic in z, 6,760 particles)
    with the wall boundary condition and the hook: the 3D wall sweeps of the
    transpose.  Fluid velocities perturbed; the gradient of <w, u> after 2
    steps in the initial positions along position directions, against
    float64 oracle central differences to 1e-5; the differences without the
    wall boundary condition differ by more than 1e-4."""
    ov = {"case": {"dim": 3, "dx": 0.05, "viscosity": 0.1},
          "solver": {"t_end": 1.0, "tvf": tvf}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("ldc", ov)
    assert solver.is_bc_trick and state["r"].shape[1] == 3
    nobc = G.setup("ldc", {**ov, "solver": {**ov["solver"], "is_bc_trick": False}})[-1]
    dt = cfg.solver.dt
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(21)
    fluid_t = (state["tag"] == 0)[:, None]
    st = dict(state)
    st["u"] = torch.where(fluid_t, state["u"] + 0.2 * torch.randn(n, 3, generator=gen).cuda(),
                          state["u"])
    st["v"] = st["u"].clone()
    if tvf:   # a transport acceleration, so the stress terms act from the first step
        st["dvdt"] = torch.where(fluid_t, torch.randn(n, 3, generator=gen).cuda(),
                                 torch.zeros_like(st["u"]))
    # the fluid sits n_walls dx up in z, past the z-periodic box (ldc.py:85): wrap it
    st["r"] = torch.remainder(st["r"], torch.as_tensor(np.asarray(box, np.float32)).cuda())
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    w = torch.randn(n, 3, generator=gen).cuda().double()
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["r"] = r0
    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, 2, tvf=tvf, checkpoint_every=1,
                              bc_fn=bc)
    ((w * u.double()).sum() + 0.0 * r.sum()).backward()
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gr).all() and np.abs(gr).max() > 0
    base = G.to_np(st)
    wn = w.cpu().numpy()

    def L(eps, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, sv, box, dt, 2, tvf=tvf, bc_fn=bc)
        return float((wn * out["u"]).sum())

    rng = np.random.default_rng(22)
    eps = 1e-6
    dirs = [gr / np.sqrt((gr ** 2).sum()), rng.standard_normal(gr.shape)]
    for trial, dr in enumerate(dirs):
        fd = (L(eps, dr) - L(-eps, dr)) / (2 * eps)
        ad = float((gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gr * dr).sum())
        print(f"3D walls tvf={tvf} trial {trial}: ad {ad:.9g} fd {fd:.9g} scale {scale:.6g}",
              flush=True)
        assert abs(ad - fd) <= 1e-5 * scale, (trial, ad, fd, scale)
        if trial == 0:
            fd_nb = (L(eps, dr, nobc) - L(-eps, dr, nobc)) / (2 * eps)
            print(f"  without the wall boundary condition: fd {fd_nb:.9g}", flush=True)
            assert abs(fd_nb - fd) > 1e-4 * abs(fd), (fd, fd_nb)


@pytest.mark.parametrize("kernel,limiter,alpha,n_steps,k", [
    ("QSK", 3.0, 0.0, 1, 1), ("QSK", 3.0, 0.0, 4, 2), ("WC2K", 3.0, 0.0, 4, 2),
    ("QSK", -1.0, 0.0, 3, 3), ("WC2K", 3.0, 0.1, 4, 2)])
def test_riemann_solver_gradient_matches_oracle_fd(kernel, limiter, alpha, n_steps, k):
    """The Riemann solver's acceleration (solver.py:316-401: the linear
    Riemann pressure P* with the dissipation limiter of solver.py:574-588, or
    the constant c_ref when eta_limiter = -1, and the viscous term) with
    density summation on the 2D TGV (dx = 0.05, velocities perturbed;
    artificial viscosity in the last case).  Gradient of
    0.5 sum(u^2) + <w, r> after n steps in the initial velocity and
    positions against float64 oracle central differences, to 1e-3 as the
    standard solver's; from two steps on, the differences with the standard
    solver differ by more than 1%.  GraphRollout equals the eager rollout."""
    ov = {"case": {"dim": 2, "dx": 0.05, "viscosity": 0.1}, "kernel": {"name": kernel},
          "solver": {"t_end": 1.0, "name": "RIE", "eta_limiter": limiter,
                     "artificial_alpha": alpha}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("tgv", ov)
    assert solver.solver == "RIE" and not solver.is_rho_evol
    std = G.setup("tgv", {**ov, "solver": {**ov["solver"], "name": "SPH"}})[-1]
    dt = cfg.solver.dt
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(31)
    st = dict(state)
    st["u"] = (state["u"].cpu() + 0.1 * torch.randn(n, 2, generator=gen)).cuda()
    st["v"] = st["u"].clone()
    st["r"] = shift(state["r"], (0.1 * 0.05 * torch.randn(n, 2, generator=gen)).cuda())
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    u0 = st["u"].clone().requires_grad_(True)
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["u"], s_in["r"] = u0, r0
    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, n_steps, checkpoint_every=k)
    (0.5 * (u.double() ** 2).sum() + (w * r.double()).sum()).backward()
    gro = adjoint.GraphRollout(st, solver, fn, shift, dt, n_steps, checkpoint_every=k)
    u2 = st["u"].clone().requires_grad_(True)
    r2 = st["r"].clone().requires_grad_(True)
    ug, rg = gro(u2, r2)
    (0.5 * (ug.double() ** 2).sum() + (w * rg.double()).sum()).backward()
    assert torch.equal(ug, u) and torch.equal(rg, r)
    assert torch.equal(u2.grad, u0.grad) and torch.equal(r2.grad, r0.grad)
    gu = u0.grad.double().cpu().numpy()
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gu).all() and np.isfinite(gr).all()
    base = G.to_np(st)
    wn = w.cpu().numpy()

    def L(eps, du, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["u"] = s["u"] + eps * du
        s["v"] = s["u"].copy()
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, sv, box, dt, n_steps)
        return 0.5 * float((out["u"] ** 2).sum()) + float((wn * out["r"]).sum())

    rng = np.random.default_rng(32)
    eps = 1e-6
    norm = np.sqrt((gu ** 2).sum() + (gr ** 2).sum())
    dirs = [(gu / norm, gr / norm)] + [(rng.standard_normal(gu.shape),
                                        rng.standard_normal(gr.shape)) for _ in range(2)]
    for trial, (du, dr) in enumerate(dirs):
        fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
        ad = float((gu * du).sum() + (gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gu * du).sum() + np.abs(gr * dr).sum())
        print(f"RIE {kernel} lim={limiter} av={alpha} n={n_steps} trial {trial}: ad {ad:.9g} "
              f"fd {fd:.9g} scale {scale:.6g}", flush=True)
        assert abs(ad - fd) <= 1e-3 * scale, (trial, ad, fd, scale)
        if trial == 0 and n_steps >= 2:
            fd_s = (L(eps, du, dr, std) - L(-eps, du, dr, std)) / (2 * eps)
            print(f"  standard solver: fd {fd_s:.9g}", flush=True)
            assert abs(fd_s - fd) > 0.01 * abs(fd), (fd, fd_s)


@pytest.mark.parametrize("case,gwbc,kernel,cont,tvf,free", [
    ("pf", True, "QSK", False, 0.0, False), ("ldc", True, "WC2K", False, 0.0, False),
    ("ldc", False, "QSK", False, 0.0, False), ("ldc", True, "WC2K", True, 0.0, False),
    ("db", True, "WC2K", True, 0.0, False), ("ldc", True, "QSK", True, 1.0, False),
    ("pf", True, "QSK", False, 0.0, True)])
def test_riemann_walls_gradient_matches_oracle_fd(case, gwbc, kernel, cont, tvf, free):
    """The Riemann solver next to walls (solver.py:345-388): a fluid particle
    sees a wall through mirrored states (u_R from the wall's velocity and
    normal, p_R = p_L + rho_L g . (-r), rho_R = R(p_R)) and, with the wall
    condition, the wall's Shepard velocity in the viscous term
    (solver.py:547-552), which passes cotangents on to the fluid around the
    wall.  The Poiseuille channel (the reference's own test runs it with this
    solver, tests/test_pf2d.py) and the lid-driven cavity, dx = 0.05, fluid
    velocities perturbed, the hook holding the walls; the gradient of <w, u>
    after 2 steps in the initial positions along position directions,
    against float64 oracle central differences to 1e-5; with the wall
    condition, the differences without it differ by more than 1e-4.  With the
    Riemann continuity (solver.py:108-164: the density of the state the step
    starts from and its pressure, the mirrored wall states with gravity), the
    cavity and the dam break as the survey's benchmark runs it (cases/db.yaml
    with the Riemann solver and Wendland C2: continuity, the wall condition,
    artificial viscosity 0.1, gravity normal to the bottom wall; dx = 0.1,
    50 warm-up steps); there the differences with density summation differ
    by more than 1e-4 as well (the dam break's tiny viscosity leaves the
    wall condition's viscous term below that)."""
    ov = {"case": {"dx": 0.05, "viscosity": 0.1} if case == "ldc" else
          {"dx": 0.1 if case == "db" else 0.05},
          "kernel": {"name": kernel},
          "solver": {"name": "RIE", "is_bc_trick": gwbc, "t_end": 1.0,
                     "density_evolution": cont, "tvf": tvf, "free_slip": free}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup(case, ov)
    assert solver.solver == "RIE" and solver.is_bc_trick == gwbc
    assert bc.kind == (2 if case == "db" else 1) and solver.is_rho_evol == cont
    plain = G.setup(case, {**ov, "solver": {**ov["solver"], "is_bc_trick": False,
                                            "free_slip": False}})[-1]
    summ = G.setup(case, {**ov, "solver": {**ov["solver"], "density_evolution": False}})[-1]
    noslip = G.setup(case, {**ov, "solver": {**ov["solver"], "free_slip": False}})[-1]
    if case == "db":
        assert solver.artificial_alpha > 0
        warm = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                       mask_self=False, format=partition.Sparse)
        nbw = warm.allocate(state["r"])
        advw = integrator.si_euler(0.0, solver.forward_wrapper(), shift, bc, None)
        for _ in range(50):
            state, nbw = advw(cfg.solver.dt, state, nbw)
    dt = cfg.solver.dt
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(41)
    fluid_t = (state["tag"] == 0)[:, None]
    st = dict(state)
    st["u"] = torch.where(fluid_t, state["u"] + 0.2 * torch.randn(n, 2, generator=gen).cuda(),
                          state["u"])
    st["v"] = st["u"].clone()
    if tvf:   # a transport acceleration, so the background-pressure term acts at once
        st["dvdt"] = torch.where(fluid_t, torch.randn(n, 2, generator=gen).cuda(),
                                 torch.zeros_like(st["u"]))
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["r"] = r0
    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, 2, tvf=tvf, checkpoint_every=1,
                              bc_fn=bc)
    ((w * u.double()).sum() + 0.0 * r.sum()).backward()
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gr).all() and np.abs(gr).max() > 0
    if solver.g_const is not None:
        gro = adjoint.GraphRollout(st, solver, fn, shift, dt, 2, tvf=tvf, checkpoint_every=1,
                                   bc_fn=bc)
        r2 = st["r"].clone().requires_grad_(True)
        ug, rg = gro(st["u"].clone(), r2)
        ((w * ug.double()).sum() + 0.0 * rg.sum()).backward()
        assert torch.equal(ug, u) and torch.equal(r2.grad, r0.grad)
    base = G.to_np(st)
    wn = w.cpu().numpy()

    def L(eps, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, sv, box, dt, 2, tvf=tvf, bc_fn=bc)
        return float((wn * out["u"]).sum())

    rng = np.random.default_rng(42)
    eps = 1e-6
    dirs = [gr / np.sqrt((gr ** 2).sum()), rng.standard_normal(gr.shape)]
    for trial, dr in enumerate(dirs):
        fd = (L(eps, dr) - L(-eps, dr)) / (2 * eps)
        ad = float((gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gr * dr).sum())
        print(f"RIE walls {case} gwbc={gwbc} cont={cont} tvf={tvf} free={free} trial "
              f"{trial}: ad {ad:.9g} fd {fd:.9g} scale {scale:.6g}", flush=True)
        assert abs(ad - fd) <= 1e-5 * scale, (trial, ad, fd, scale)
        if trial == 0 and free:
            fd_n = (L(eps, dr, noslip) - L(-eps, dr, noslip)) / (2 * eps)
            print(f"  no-slip walls: fd {fd_n:.9g}", flush=True)
            assert abs(fd_n - fd) > 1e-4 * abs(fd), (fd, fd_n)
        if trial == 0 and gwbc and case != "db" and not free:
            # (the dam break's viscosity, 5e-5, leaves the wall condition's
            # viscous term too small to see over two steps)
            fd_p = (L(eps, dr, plain) - L(-eps, dr, plain)) / (2 * eps)
            print(f"  without the wall condition: fd {fd_p:.9g}", flush=True)
            assert abs(fd_p - fd) > 1e-4 * abs(fd), (fd, fd_p)
        if trial == 0 and cont:
            fd_s = (L(eps, dr, summ) - L(-eps, dr, summ)) / (2 * eps)
            print(f"  with density summation: fd {fd_s:.9g}", flush=True)
            assert abs(fd_s - fd) > 1e-4 * abs(fd), (fd, fd_s)


@pytest.mark.parametrize("case,solver_name,gwbc", [("ldc", "SPH", False), ("ldc", "SPH", True),
                                                 ("ldc", "RIE", True), ("db", "RIE", True)])
def test_density_renormalisation_gradient_matches_oracle_fd(case, solver_name, gwbc):
    """Density renormalisation after continuity (solver.py:167-173:
    rho = sum m W / min(sum (m/rho) W, 1) over all pairs): with the standard
    solver on the cavity with plain held walls and with the wall condition,
    and with the Riemann solver on the cavity and the dam break (dx = 0.1,
    50 warm-up steps), fluid positions jittered by 0.1 dx (on the lattice the
    Shepard sums sit at the clamp, a kink that differences straddle).  The
    gradient of <w, u> after 2 steps in the initial
    positions along position directions, against float64 oracle central
    differences to 1e-5; the differences without renormalisation differ by
    more than 1e-4."""
    ov = {"case": {"dx": 0.05, "viscosity": 0.1} if case == "ldc" else {"dx": 0.1},
          "kernel": {"name": "WC2K" if solver_name == "RIE" else "QSK"},
          "solver": {"name": solver_name, "is_bc_trick": gwbc, "t_end": 1.0,
                     "density_evolution": True, "density_renormalize": True}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup(case, ov)
    assert solver.is_rho_evol and solver.is_rho_renorm
    plain = G.setup(case, {**ov, "solver": {**ov["solver"], "density_renormalize": False}})[-1]
    if case == "db":
        warm = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                       mask_self=False, format=partition.Sparse)
        nbw = warm.allocate(state["r"])
        advw = integrator.si_euler(0.0, solver.forward_wrapper(), shift, bc, None)
        for _ in range(50):
            state, nbw = advw(cfg.solver.dt, state, nbw)
    dt = cfg.solver.dt
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(51)
    fluid_t = (state["tag"] == 0)[:, None]
    st = dict(state)
    st["u"] = torch.where(fluid_t, state["u"] + 0.2 * torch.randn(n, 2, generator=gen).cuda(),
                          state["u"])
    st["v"] = st["u"].clone()
    # jittered fluid: on the lattice the Shepard sums sit at the clamp min(D, 1),
    # where the renormalisation has a kink that differences straddle
    dxc = float(cfg.case.dx)
    st["r"] = torch.where(fluid_t, shift(st["r"], (0.1 * dxc * torch.randn(n, 2, generator=gen)
                                                   ).cuda()), st["r"])
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    r0 = st["r"].clone().requires_grad_(True)
    s_in = dict(st)
    s_in["r"] = r0
    u, r, _ = adjoint.rollout(s_in, solver, fn, shift, dt, 2, checkpoint_every=1, bc_fn=bc)
    ((w * u.double()).sum() + 0.0 * r.sum()).backward()
    gr = r0.grad.double().cpu().numpy()
    assert np.isfinite(gr).all() and np.abs(gr).max() > 0
    gro = adjoint.GraphRollout(st, solver, fn, shift, dt, 2, checkpoint_every=1, bc_fn=bc)
    r2 = st["r"].clone().requires_grad_(True)
    ug, rg = gro(st["u"].clone(), r2)
    ((w * ug.double()).sum() + 0.0 * rg.sum()).backward()
    assert torch.equal(ug, u) and torch.equal(r2.grad, r0.grad)
    base = G.to_np(st)
    wn = w.cpu().numpy()

    def L(eps, dr, sv=solver):
        s = G.to_f64(dict(base))
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, sv, box, dt, 2, bc_fn=bc)
        return float((wn * out["u"]).sum())

    rng = np.random.default_rng(52)
    eps = 1e-6
    dirs = [gr / np.sqrt((gr ** 2).sum()), rng.standard_normal(gr.shape)]
    for trial, dr in enumerate(dirs):
        fd = (L(eps, dr) - L(-eps, dr)) / (2 * eps)
        ad = float((gr * dr).sum())
        scale = abs(fd) if trial == 0 else float(np.abs(gr * dr).sum())
        print(f"renorm {case} {solver_name} gwbc={gwbc} trial {trial}: ad {ad:.9g} fd {fd:.9g} "
              f"scale {scale:.6g}", flush=True)
        assert abs(ad - fd) <= 1e-5 * scale, (trial, ad, fd, scale)
        if trial == 0:
            fd_p = (L(eps, dr, plain) - L(-eps, dr, plain)) / (2 * eps)
            print(f"  without renormalisation: fd {fd_p:.9g}", flush=True)
            assert abs(fd_p - fd) > 1e-4 * abs(fd), (fd, fd_p)


def test_heat_conduction_leaves_the_velocity_and_position_gradient():
    """With heat conduction on (solver.py:832-850) the temperature evolves
    beside the flow but enters neither the EOS nor the forces: the rollout's
    u and r, and so their gradient, are those of the same run without it.
    2D TGV (dx = 0.05) with a nonuniform temperature; the gradient of
    0.5 sum(u^2) + <w, r> after 3 steps equals the one without heat to float32
    rounding and matches float64 oracle central differences of the run with
    heat to 1e-3."""
    ov = {"case": {"dim": 2, "dx": 0.05, "viscosity": 0.1},
          "solver": {"t_end": 1.0, "heat_conduction": True}}
    cfg, box, state, g, bc, eos, disp, shift, solver = G.setup("tgv", ov)
    plain = G.setup("tgv", {**ov, "solver": {"t_end": 1.0}})[-1]
    assert solver.is_heat_conduction and not plain.is_heat_conduction
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(4)
    st = dict(state)
    st["T"] = (1.0 + 0.2 * torch.rand(n, generator=gen)).cuda()
    st["kappa"] = torch.full_like(st["T"], 0.5)    # the case's defaults are 0
    st["Cp"] = torch.full_like(st["T"], 2.0)
    st["u"] = (state["u"].cpu() + 0.05 * torch.randn(n, 2, generator=gen)).cuda()
    st["v"] = st["u"].clone()
    dt = cfg.solver.dt
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    w = torch.randn(n, 2, generator=gen).cuda().double()
    grads = []
    for sv in (solver, plain):
        u0 = st["u"].clone().requires_grad_(True)
        r0 = st["r"].clone().requires_grad_(True)
        s_in = dict(st)
        s_in["u"], s_in["r"] = u0, r0
        u, r, _ = adjoint.rollout(s_in, sv, fn, shift, dt, 3, checkpoint_every=2)
        (0.5 * (u.double() ** 2).sum() + (w * r.double()).sum()).backward()
        grads.append((u0.grad.double().cpu().numpy(), r0.grad.double().cpu().numpy()))
    (gu, gr), (gu0, gr0) = grads
    scale = np.abs(gu).max() + np.abs(gr).max()
    assert np.abs(gu - gu0).max() <= 1e-6 * scale and np.abs(gr - gr0).max() <= 1e-6 * scale
    base = G.to_np(st)
    wn = w.cpu().numpy()

    def L(eps, du, dr):
        s = G.to_f64(dict(base))
        s["u"] = s["u"] + eps * du
        s["v"] = s["u"].copy()
        s["r"] = np.mod(s["r"] + eps * dr, box)
        out = _oracle_rollout(s, solver, box, dt, 3)
        assert np.ptp(out["T"]) > 0.1 and np.isfinite(out["T"]).all()   # a live temperature
        return 0.5 * float((out["u"] ** 2).sum()) + float((wn * out["r"]).sum())

    eps = 1e-6
    norm = np.sqrt((gu ** 2).sum() + (gr ** 2).sum())
    du, dr = gu / norm, gr / norm
    fd = (L(eps, du, dr) - L(-eps, du, dr)) / (2 * eps)
    ad = float((gu * du).sum() + (gr * dr).sum())
    assert abs(ad - fd) <= 1e-3 * abs(fd), (ad, fd)


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
    # delta-SPH's continuity (its density diffusion) is not differentiated
    delta = G.setup("tgv", {"case": {"dim": 2, "dx": 0.05},
                            "solver": {"name": "DELTA", "density_evolution": True}})[-1]
    with pytest.raises(NotImplementedError, match="density summation only"):
        adjoint.rollout(st, delta, fn, shift, cfg.solver.dt, 2)
    skin_fn = partition.neighbor_list(solver.displacement_fn, box,
                                      r_cutoff=solver._kernel_fn.cutoff, mask_self=False,
                                      format=partition.Sparse)
    with pytest.raises(ValueError):
        adjoint.rollout(st, solver, skin_fn, shift, cfg.solver.dt, 2)
    # a position-dependent body force: not under graph replay, not with continuity
    from jax_sph_amd import case_setup as cs
    ov = {"case": {"dim": 2, "dx": 0.05}}
    c2, _, _, _, _, eos, disp, _, _ = G.setup("tgv", ov)
    forced = cs.make_solver(c2, disp, eos, _SmoothForce(1.0))
    with pytest.raises(NotImplementedError, match="constant body force"):
        adjoint.GraphRollout(st, forced, fn, shift, cfg.solver.dt, 2, checkpoint_every=1)


def _c4_fixture():
    import json
    import os
    gdir = os.path.join(os.path.dirname(__file__), "golden")
    fx = json.load(open(os.path.join(gdir, "c4_fd.json")))
    target = np.load(os.path.join(gdir, "c4_target_u.npy"), allow_pickle=False)
    return fx, target


def test_c4_gradient_matches_float64_finite_differences():
    """2D TGV inverse problem at full size (10,000 particles, 100 steps,
    checkpoint every 10) against 32 central differences (eps = 1e-3) of the
    float64 oracle rollout, committed in tests/golden/c4_fd.json (made by
    tests/golden/make_c4_fd.py):

    * the loss the differences were taken of, MSE(u_final, u_target) with the
      float64 target rollout (tests/golden/c4_target_u.npy), to 1e-3 relative
      (vector norm over the 32 directions);
    * the linear functional <w, u_final>, whose gradient does not involve the
      rollout's residual, to 1e-3 relative: this pins the transposed steps
      themselves at full size.

    The float64 differences are converged (eps = 5e-4, 1e-3 and 2e-3 agree to
    1e-5).  The graph-replayed rollout must give the eager gradient bit for
    bit."""
    from jax_sph_amd import case_setup as cs
    fx, target_np = _c4_fixture()
    cfg = cs.load_config("tgv", {"case": {"dim": 2, "dx": 0.01}, "solver": {"tvf": 0.0}})
    (cfg, box, state, g, bc, nw, eos, key, disp, shift) = cs.load_case("tgv")(cfg).initialize()
    solver = cs.make_solver(cfg, disp, eos, g)
    assert repr(cfg.solver.dt) in fx["problem"]
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff, mask_self=False,
                                 format=partition.Sparse, skin=0.0)
    n = state["r"].shape[0]
    assert target_np.shape == (n, 2)
    target = torch.from_numpy(target_np).cuda()
    gen = torch.Generator().manual_seed(fx["seed"])
    u0 = (state["u"].cpu() + 0.05 * torch.randn(n, 2, generator=gen)).cuda()
    dirs = [torch.randn(n, 2, generator=gen).cuda().double() for _ in range(len(fx["fd"]))]
    w = torch.randn(n, 2, generator=gen).cuda().double()

    def grad_of(loss_fn):
        uu = u0.clone().requires_grad_(True)
        s = dict(state)
        s["u"] = uu
        u, _, _ = adjoint.rollout(s, solver, fn, shift, cfg.solver.dt, fx["steps"],
                                  checkpoint_every=10)
        loss_fn(u.double()).backward()
        return uu.grad

    def rel(grad, fd):
        ad = np.array([float((grad.double() * d).sum()) for d in dirs])
        fd = np.array(fd)
        return np.linalg.norm(ad - fd) / np.linalg.norm(fd), ad, fd

    g_mse = grad_of(lambda u: ((u - target) ** 2).mean())
    vec, ad, fd = rel(g_mse, fx["fd"])
    print(f"C4 MSE gradient: relative error {vec:.3e} over {len(fd)} directions")
    assert vec <= 1e-3, (vec, ad[:4], fd[:4])
    g_lin = grad_of(lambda u: (w * u).sum())
    vec_l, ad, fd = rel(g_lin, fx["fd_linear"])
    print(f"C4 linear-functional gradient: relative error {vec_l:.3e}")
    assert vec_l <= 1e-3, (vec_l, ad[:4], fd[:4])
    # and the graph-replayed rollout gives the same gradient
    gr = adjoint.GraphRollout(state, solver, fn, shift, cfg.solver.dt, fx["steps"],
                              checkpoint_every=10)
    u2 = u0.clone().requires_grad_(True)
    ug, _ = gr(u2, state["r"])
    ((ug.double() - target) ** 2).mean().backward()
    assert torch.equal(u2.grad, g_mse)


@pytest.mark.parametrize("sets", [None, 1, 2])
def test_graph_rollout_equals_eager_rollout(sets):
    """GraphRollout (both passes as HIP graphs, one workspace per step of a
    segment) gives the eager rollout's values and gradients bit for bit, also
    when called twice: with a workspace set per segment (the forward keeps
    every step, the backward only applies the transposes; None picks that at
    this size) and with fewer sets than segments (the backward recomputes).
    Two forwards before the first backward make the kept sets stale, and the
    backward of the first then recomputes its segments."""
    cfg, box, st, solver, fn, shift, gen = _setup()
    n = st["r"].shape[0]
    w = torch.randn(n, 2, generator=gen).cuda()
    u0 = st["u"].clone().requires_grad_(True)
    s = dict(st)
    s["u"] = u0
    u, r, _ = adjoint.rollout(s, solver, fn, shift, cfg.solver.dt, 12, checkpoint_every=4)
    ((w * u).sum() + (w * r).sum()).backward()
    g_eager = u0.grad.clone()
    gr = adjoint.GraphRollout(st, solver, fn, shift, cfg.solver.dt, 12, checkpoint_every=4,
                              workspace_sets=sets)
    assert gr.keep_all == (sets is None)
    for _ in range(2):
        uu = st["u"].clone().requires_grad_(True)
        ug, rg = gr(uu, st["r"])
        ((w * ug).sum() + (w * rg).sum()).backward()
        torch.cuda.synchronize()
        assert torch.equal(ug, u) and torch.equal(rg, r)
        assert torch.equal(uu.grad, g_eager)
    # a second forward (from another start) between the first and its backward
    u1 = st["u"].clone().requires_grad_(True)
    ug1, rg1 = gr(u1, st["r"])
    u2 = (st["u"] * 0.5).requires_grad_(True)
    gr(u2, st["r"])
    ((w * ug1).sum() + (w * rg1).sum()).backward()
    torch.cuda.synchronize()
    assert torch.equal(u1.grad, g_eager)

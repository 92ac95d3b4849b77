"""Throughput of the SPH step on one MI355X: 3D Taylor-Green vortex, 262,144 particles.

Workload (the 3D TGV configuration): periodic cube [0, 2 pi]^3, 64^3 particles on
a lattice, float32, Quintic kernel h = dx (cutoff 3h), transport-velocity SPH
(tvf = 1), density summation, Tait EOS c0 = 10, Re = 1600, dt from the CFL
condition (2.2312e-3).  A step is one fused si_euler step: kick + drift +
periodic wrap, cell binning, neighbour search, density + EOS, forces.

Prints one JSON line.  ``value`` = particle-steps/s over the timed steps (inputs
resident in HBM).  ``roofline`` = the dominant kernel's algorithmic bytes per
launch / its mean device duration (HIP events on the library's stream).
``cpu_baseline`` = the OpenMP C restatement of the step (oracle/sph_cpu.c, a
CPU restatement of jax-sph) on the same workload, on this host's threads, for
about 15 s.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "jax-sph_amd"))

HBM_PEAK_GBS = 8000.0
FP32_VALU_TFLOPS = 157.3

# algorithmic (compulsory) bytes per particle and launch, 3D float32:
# force: reads r, u, v, rho, p, mass, eta, tag (14 x 4 B), writes dudt, dvdt (6 x 4 B)
# density: reads r, mass, tag (5 x 4 B), writes rho, p (2 x 4 B)
# nl_build: reads r (12 B), writes one int32 per pair (4 B x pairs / N)
ALG_BYTES_3D = {"force": 80.0, "density": 28.0}
# algorithmic float32 operations per pair inside the cutoff (add, mul, div,
# sqrt = 1, FMA = 2), counted on the reference formulas, 3D, QSK:
# density (solver.py:176-178, kernel.py:51-72): displacement 3, d^2 5, sqrt 1,
#   q 1, W(q) 20 (three clamped fifth powers, weights, sigma), sum 1 = 31;
# force (solver.py:199-256, tvf = 1, one viscosity): displacement 3, d^2 5,
#   sqrt 1, q 1, W'(q) 17, 1/(d + EPS) 2, weighted volume 4, prefactor 2,
#   eta_ij 5, p_ij 5, A_i r 9, A_j r 9, (A_i + A_j) r / 2 6, -p_ij r 3,
#   eta_ij u_ij 6, sum of terms 6, times prefactor 3, accumulate 3,
#   background-pressure acceleration 4 + accumulate 3 = 97.
ALG_FLOPS_PER_PAIR_3D = {"force": 97.0, "density": 31.0}
COUNTERS = os.path.join(ROOT, "profiles", "counters_tgv3d_64cubed_sweeps.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n-side", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the 2D TGV forward + adjoint leg")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-member batched ensemble leg")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the ensemble leg and print its JSON")
    ap.add_argument("--cpu-sample-side", type=int, default=64)
    ap.add_argument("--skin", default="auto",
                    help="Verlet skin of the fused step's pair lists ('auto' = 0.08 cutoff, "
                         "0 = exact lists rebuilt every step)")
    ap.add_argument("--xcd-remap", type=int, default=1,
                    help="1: XCD-contiguous block order in the pair build and sweeps")
    ap.add_argument("--scatter-probe", action="store_true",
                    help="only time the j-side scatter-adds that half-pair (Newton's third "
                         "law) sweeps would need at --n-side, and print them")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (5 times), to "
                         "scale hardware byte counters; prints nothing else")
    return ap.parse_args()


def copy_calibration(mb):
    import torch
    n = mb * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mb, "launches": 5}))


def scatter_probe(n_side, reps=5):
    """Cost of the j-side updates of half-pair sweeps on the 3D TGV lattice:
    each unordered pair inside the cutoff is evaluated once and its result is
    added to both particles, so the pair's second particle receives a
    scatter-add (6 floats for du/dt and dv/dt, 1 for a density sum).  The
    pairs are the exact neighbour list at n_side^3 in its sender-major order,
    the order a sweep visits them; torch's index_add_ is the scatter (float
    atomics on the device)."""
    import torch
    from jax_sph_amd import case_setup as cs
    from jax_sph_amd import partition
    cfg = cs.load_config("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / n_side}})
    (cfg, box, state, g_ext_fn, bc_fn, nw_fn, eos, key, disp, shift) = \
        cs.load_case("tgv")(cfg).initialize()
    solver = cs.make_solver(cfg, disp, eos, g_ext_fn)
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=True,
                                 format=partition.Sparse, skin=0.0)
    nb = fn.allocate(state["r"])
    n = state["r"].shape[0]
    recv, send = nb.idx[0].long(), nb.idx[1].long()
    half = (recv < n) & (recv > send)
    recv = recv[half].contiguous()
    m = int(recv.numel())
    out = {"n": n, "half_pairs": m}
    for name, width in (("force", 6), ("density", 1)):
        vals = torch.ones((m, width), dtype=torch.float32, device="cuda")
        acc = torch.zeros((n, width), dtype=torch.float32, device="cuda")
        acc.index_add_(0, recv, vals)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            acc.index_add_(0, recv, vals)
        ev[1].record()
        torch.cuda.synchronize()
        out[f"{name}_scatter_us"] = ev[0].elapsed_time(ev[1]) * 1e3 / reps
        out[f"{name}_float_adds"] = m * width
    print(json.dumps({"scatter_probe": out}), flush=True)


def fwd_adjoint_leg(steps=100, every=10, seed=0):
    """2D Taylor-Green inverse problem: dx = 0.01 (10,000 particles), tvf = 0,
    a rollout of `steps` si_euler steps; loss = MSE of the final velocities
    against the rollout from the analytic initial field; gradient with respect
    to the initial velocity (analytic + N(0, 0.05^2)) through the hand-written
    adjoint with a checkpoint every `every` steps, both passes replayed as HIP
    graphs.  Reports t(forward with checkpoints + adjoint) / t(forward), the
    forward being the same rollout as graph replays of the fused step, and the
    gradient error against the float64 oracle's central differences committed
    in tests/golden/c4_fd.json (same u0 and directions)."""
    import torch
    from jax_sph_amd import adjoint, integrator, partition
    from jax_sph_amd import case_setup as cs
    from jax_sph_amd.graph import StepGraph
    cfg = cs.load_config("tgv", {"case": {"dim": 2, "dx": 0.01}, "solver": {"tvf": 0.0}})
    (cfg, box, state, g_ext_fn, bc_fn, nw_fn, eos, key, disp, shift) = \
        cs.load_case("tgv")(cfg).initialize()
    solver = cs.make_solver(cfg, disp, eos, g_ext_fn)
    dt = cfg.solver.dt
    cut = solver._kernel_fn.cutoff
    fast_fn = partition.neighbor_list(disp, box, r_cutoff=cut, mask_self=False,
                                      format=partition.Sparse)
    exact_fn = partition.neighbor_list(disp, box, r_cutoff=cut, mask_self=False,
                                       format=partition.Sparse, skin=0.0)
    adv = integrator.si_euler(0.0, solver.forward_wrapper(), shift, bc_fn, nw_fn)
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(seed)
    u0 = (state["u"].cpu() + 0.05 * torch.randn(n, 2, generator=gen)).cuda()

    # plain forward: `every` fused steps per graph replay
    sg = StepGraph(adv, state, fast_fn.allocate(state["r"]), dt, steps_per_replay=every)
    base = {k: v.clone() for k, v in sg.state.items()}

    def plain(u_init):
        for k in sg.state:
            sg.state[k].copy_(base[k])
        sg.state["u"].copy_(u_init)
        sg.state["v"].copy_(u_init)
        sg.replay(steps // every)
        return sg.state["u"].clone()

    # the float64 oracle's target rollout (tests/golden/c4_target_u.npy), the
    # target the committed finite differences were taken with; without the
    # fixture, the float32 rollout of the analytic field
    gdir = os.path.join(ROOT, "tests", "golden")
    tpath = os.path.join(gdir, "c4_target_u.npy")
    if os.path.exists(tpath):
        u_tgt = torch.from_numpy(np.load(tpath, allow_pickle=False)).cuda()
        target_kind = "float64 oracle rollout of the analytic field (tests/golden/c4_target_u.npy)"
    else:
        u_tgt = plain(state["u"]).double()
        target_kind = "float32 rollout of the analytic field"
    plain(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain(u0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    assert not sg.did_buffer_overflow

    gr = adjoint.GraphRollout(state, solver, exact_fn, shift, dt, steps, checkpoint_every=every)

    def fwd_adj(loss_fn=lambda u: ((u - u_tgt) ** 2).mean()):
        uu = u0.clone().requires_grad_(True)
        u_n, _ = gr(uu, state["r"])
        loss = loss_fn(u_n.double())
        loss.backward()
        return loss, uu.grad

    fwd_adj()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, grad = fwd_adj()
    torch.cuda.synchronize()
    t_fa = time.perf_counter() - t0
    out = {"config": f"2D TGV dx=0.01 ({n} particles), tvf=0, {steps} steps, checkpoint "
                     f"every {every}, loss = MSE(u_final, target rollout)",
           "t_forward_s": t_fwd, "t_forward_plus_adjoint_s": t_fa,
           "cost_ratio": t_fa / t_fwd, "loss": float(loss), "target": target_kind}
    fx_path = os.path.join(gdir, "c4_fd.json")
    if os.path.exists(fx_path) and os.path.exists(tpath):
        fx = json.load(open(fx_path))
        dirs = [torch.randn(n, 2, generator=gen).cuda().double() for _ in range(len(fx["fd"]))]
        w = torch.randn(n, 2, generator=gen).cuda().double()
        _, g_lin = fwd_adj(lambda u: (w * u).sum())

        def check(g, fd):
            ad = np.array([float((g.double() * d).sum()) for d in dirs])
            fd = np.array(fd)
            rel = np.abs(ad - fd) / np.abs(fd)
            return {"rel_err_vector": float(np.linalg.norm(ad - fd) / np.linalg.norm(fd)),
                    "rel_err_median": float(np.median(rel)), "rel_err_max": float(np.max(rel))}

        out["grad_check"] = {"against": "float64 oracle central differences, eps=1e-3, "
                                        "tests/golden/c4_fd.json", "directions": len(dirs),
                             **check(grad, fx["fd"]),
                             "linear_functional": check(g_lin, fx["fd_linear"])}
    return out


def fwd_adjoint_3d_leg(n_side=64, steps=100, every=10, seed=0):
    """The 3D headline configuration differentiated: 3D TGV n_side^3 with the
    transport-velocity formulation (tvf = 1), a rollout of `steps` si_euler
    steps with a checkpoint every `every` steps, both passes replayed as HIP
    graphs (adjoint.GraphRollout on exact pair lists), the gradient of
    <w, u_final> + <w, r_final> with respect to the initial velocity.  Reports
    t(forward with checkpoints + adjoint) against the plain forward of the
    same rollout: graph-replayed fused steps on the skin lists the benchmark
    times, and on exact lists rebuilt every step as the adjoint's forward
    runs."""
    import torch
    from jax_sph_amd import adjoint, integrator, partition
    from jax_sph_amd import case_setup as cs
    from jax_sph_amd.graph import StepGraph
    cfg = cs.load_config("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / n_side,
                                          "viscosity": 1 / 1600},
                                 "solver": {"tvf": 1.0, "t_end": 1.0}})
    (cfg, box, state, g_ext_fn, bc_fn, nw_fn, eos, key, disp, shift) = \
        cs.load_case("tgv")(cfg).initialize()
    solver = cs.make_solver(cfg, disp, eos, g_ext_fn)
    dt = cfg.solver.dt
    cut = solver._kernel_fn.cutoff
    n = state["r"].shape[0]
    gen = torch.Generator().manual_seed(seed)
    u0 = (state["u"].cpu() + 0.05 * torch.randn(n, 3, generator=gen)).cuda()
    w = torch.randn(n, 3, generator=gen).cuda()
    adv = integrator.si_euler(1.0, solver.forward_wrapper(), shift, bc_fn, nw_fn)
    timings = {}
    for name, skin in (("t_forward_skin_s", "auto"), ("t_forward_exact_s", 0.0)):
        fn = partition.neighbor_list(disp, box, r_cutoff=cut, capacity_multiplier=1.25,
                                     mask_self=False, format=partition.Sparse, skin=skin)
        sg = StepGraph(adv, state, fn.allocate(state["r"]), dt, steps_per_replay=every)
        base = {k: v.clone() for k, v in sg.state.items()}
        for rep in range(2):   # warm-up, then timed
            for k in sg.state:
                sg.state[k].copy_(base[k])
            sg.state["u"].copy_(u0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sg.replay(steps // every)
            torch.cuda.synchronize()
            timings[name] = time.perf_counter() - t0
        assert not sg.did_buffer_overflow
        del sg, base
    exact_fn = partition.neighbor_list(disp, box, r_cutoff=cut, capacity_multiplier=1.25,
                                       mask_self=False, format=partition.Sparse, skin=0.0)
    t0 = time.perf_counter()
    gr = adjoint.GraphRollout(state, solver, exact_fn, shift, dt, steps, tvf=1.0,
                              checkpoint_every=every)
    setup_s = time.perf_counter() - t0

    def fwd_adj():
        uu = u0.clone().requires_grad_(True)
        u_n, r_n = gr(uu, state["r"])
        ((w * u_n.double()).sum() + (w * r_n.double()).sum()).backward()
        return uu.grad

    fwd_adj()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grad = fwd_adj()
    torch.cuda.synchronize()
    t_fa = time.perf_counter() - t0
    assert bool(torch.isfinite(grad).all())
    return {"config": f"3D TGV {n_side}^3 ({n} particles), tvf=1, Re=1600, {steps} steps, "
                      f"checkpoint every {every}, gradient of <w,u_final>+<w,r_final> in u0",
            "t_forward_skin_s": timings["t_forward_skin_s"],
            "t_forward_exact_s": timings["t_forward_exact_s"],
            "t_forward_plus_adjoint_s": t_fa, "graph_setup_s": setup_s,
            "workspace_sets": gr.n_sets, "segments_kept_in_place": gr.keep_all,
            "workspace_gb": sum(nb.workspace.bytes for st_ in gr.sets for nb in st_) / 1e9,
            "cost_ratio": t_fa / timings["t_forward_skin_s"],
            "cost_ratio_vs_exact_forward": t_fa / timings["t_forward_exact_s"],
            "particle_steps_per_s_fwd_adjoint": n * steps / t_fa}


ENSEMBLE_DX = (0.025, 0.02, 0.0125, 0.01, 0.005)


def ensemble_configs(members=64, seed=0):
    """The batched-ensemble workload: the first half 2D Taylor-Green vortices
    (periodic unit square, standard SPH), the second half lid-driven cavities
    (lid u = 1, Re = 100, 3 wall layers, generalised wall BC); each dx drawn
    from ENSEMBLE_DX with a seeded generator; dt from each member's CFL
    condition (1,600 to 42,436 particles per member)."""
    from jax_sph_amd import case_setup as cs
    dxs = np.random.default_rng(seed).choice(ENSEMBLE_DX, size=members)
    cfgs = []
    for k, dx in enumerate(dxs):
        if k < members // 2:
            cfgs.append(cs.load_config("tgv", {"case": {"dim": 2, "dx": float(dx)}}))
        else:
            cfgs.append(cs.load_config("ldc", {"case": {"dx": float(dx)},
                                               "solver": {"dt": None}}))
    return cfgs, [float(x) for x in dxs]


def ensemble_leg(steps=1000, members=64, seed=0, graph_steps=20):
    """64 uneven members in one launch per step phase (graph-replayed; each
    group's skin width and the block layout autotuned) against the same 64 runs
    one after the other, each on its own fused step with the same skin, eager
    and graph-replayed."""
    import torch
    from jax_sph_amd.ensemble import Ensemble, setup_member
    from jax_sph_amd.graph import StepGraph

    reps = max(1, steps // graph_steps)
    steps = reps * graph_steps
    cfgs, dxs = ensemble_configs(members, seed)
    t0 = time.perf_counter()
    ens = Ensemble.from_configs(cfgs)
    setup_s = time.perf_counter() - t0
    del ens
    # launch shapes: each group's Verlet-skin width, then the block layout
    t0 = time.perf_counter()
    ens, tune = Ensemble.tuned_from_configs(cfgs, steps=graph_steps, reps=2)
    tune["seconds"] = time.perf_counter() - t0
    factors = ens.skin_factors
    info = ens.info()
    ens.capture(graph_steps)
    ens.replay(1)  # warm-up; the timed replays continue the run
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.replay(reps)
    torch.cuda.synchronize()
    batched = time.perf_counter() - t0
    overflow_batched = ens.overflowed()
    del ens
    seq_eager = seq_graph = 0.0
    overflow_seq = 0
    for cfg, f in zip(ensemble_configs(members, seed)[0], factors):
        adv, st, nb, dt = setup_member(cfg, skin_factor=f)  # the member's own skin
        st, nb = adv(dt, st, nb)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st, nb = adv(dt, st, nb)
        torch.cuda.synchronize()
        seq_eager += time.perf_counter() - t0
        overflow_seq |= int(nb.error.code)
        g = StepGraph(adv, st, nb, dt, steps_per_replay=graph_steps)
        g.replay(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.replay(reps)
        torch.cuda.synchronize()
        seq_graph += time.perf_counter() - t0
        overflow_seq |= int(g.did_buffer_overflow)
        del g, st, nb, adv
    return {
        "workload": f"{members} 2D members: {members // 2} Taylor-Green vortices + "
                    f"{members - members // 2} lid-driven cavities (Re 100, 3 wall layers), "
                    f"dx drawn from {list(ENSEMBLE_DX)} (seed {seed}), CFL dt, QSK h=dx",
        "members": info["members"], "particles": info["particles"],
        "sweep_groups": info["groups"], "steps": steps, "setup_s": setup_s,
        "batched_ms_per_step": batched / steps * 1e3,
        "sequential_eager_ms_per_step": seq_eager / steps * 1e3,
        "sequential_graph_ms_per_step": seq_graph / steps * 1e3,
        "speedup_vs_sequential_eager": seq_eager / batched,
        "speedup_vs_sequential_graph": seq_graph / batched,
        "member_steps_per_s_batched": members * steps / batched,
        "particle_steps_per_s_batched": info["particles"] * steps / batched,
        "autotune": tune, "overflowed_members": overflow_batched,
        "sequential_error_code": overflow_seq, "dx": dxs,
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(n_side, seconds=15.0):
    """The same 3D TGV workload (n_side^3 particles, tvf = 1, QSK, Re 1600,
    CFL dt) stepped by the OpenMP C restatement of the jax-sph step
    (oracle/sph_cpu.c, checked against the numpy oracle in
    tests/test_cpu_port.py) on every host thread the run may use, for about
    `seconds` of wall time."""
    sys.path.insert(0, ROOT)
    from oracle import cpu_port
    from oracle import sph_oracle as so
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    dx = 2 * np.pi / n_side
    box = np.full(3, 2 * np.pi)
    i = np.arange(n_side)
    g = np.meshgrid(i, i, i, indexing="xy")
    idx = np.vstack([x.ravel() for x in g]).T
    r = ((idx + np.float32(0.5)).astype(np.float32) * np.float32(dx)).astype(np.float32)
    N = len(r)
    x, y, z = r[:, 0], r[:, 1], r[:, 2]
    u = np.zeros_like(r)
    u[:, 0] = np.sin(x) * np.cos(y) * np.cos(z)
    u[:, 1] = -np.cos(x) * np.sin(y) * np.cos(z)
    f = np.float32
    st = {"r": r, "u": u, "v": u.copy(), "dudt": np.zeros_like(r), "dvdt": np.zeros_like(r),
          "rho": np.ones(N, f), "p": np.zeros(N, f), "mass": np.full(N, f(dx ** 3), f),
          "eta": np.full(N, f(1 / 1600), f)}
    kernel = so.Kernel("QSK", dx, 3)
    eos = so.TaitEoS(100.0, 1.0, 0.0, 1.0)
    dt = 0.25 * dx / 11.0
    step = cpu_port.CPUStep(box, kernel, eos, dt, 1.0, threads=threads)
    step(st)  # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        step(st)
        steps += 1
        if time.perf_counter() - t0 > seconds:
            break
    el = time.perf_counter() - t0
    return {"value": N * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"OpenMP C restatement of the jax-sph step (oracle/sph_cpu.c; a CPU "
                      f"restatement of jax-sph, not jax-sph), 3D TGV {n_side}^3 = {N} "
                      f"particles, tvf=1, {steps} si_euler steps in {el:.1f} s on {threads} "
                      f"threads, {_cpu_model()}, nproc={os.cpu_count()}"}


def main():
    a = parse()
    if a.copy_calibration_mb:
        copy_calibration(a.copy_calibration_mb)
        return
    if a.scatter_probe:
        scatter_probe(a.n_side)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not "
                                   "part of it (--gpus must be 1)"}))
        sys.exit(2)
    import torch
    import jax_sph_amd as js
    from jax_sph_amd import case_setup as cs
    from jax_sph_amd import integrator, partition

    if a.ensemble_only:
        print(json.dumps({"ensemble": ensemble_leg(steps=a.ensemble_steps)}), flush=True)
        return
    n_side = a.n_side
    cfg = cs.load_config("tgv", {"case": {"dim": 3, "dx": 2 * np.pi / n_side,
                                          "viscosity": 1 / 1600},
                                 "solver": {"tvf": 1.0, "t_end": 1.0}})
    (cfg, box, state, g_ext_fn, bc_fn, nw_fn, eos, key, disp, shift) = \
        cs.load_case("tgv")(cfg).initialize()
    solver = cs.make_solver(cfg, disp, eos, g_ext_fn)
    skin = a.skin if a.skin == "auto" else float(a.skin)
    fn = partition.neighbor_list(disp, box, r_cutoff=solver._kernel_fn.cutoff,
                                 capacity_multiplier=1.25, mask_self=False,
                                 format=partition.Sparse, skin=skin)
    nb = fn.allocate(state["r"])
    nb.workspace.tune("xcd_remap", a.xcd_remap)
    adv = integrator.si_euler(cfg.solver.tvf, solver.forward_wrapper(), shift, bc_fn, nw_fn)
    assert adv._sph_fusable(nb) is not None, "bench must run the fused step"
    N = state["r"].shape[0]
    dt = cfg.solver.dt
    ws = nb.workspace
    st = state
    for _ in range(a.warmup):
        st, nb = adv(dt, st, nb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        st, nb = adv(dt, st, nb)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    overflow = int(nb.error.code)
    value = N * a.steps / el
    # as many steps again, continuing the run, with a HIP event pair around
    # every library kernel (on the library's stream): mean device durations
    ws.prof_reset()
    ws.prof_enable(True)
    t1 = time.perf_counter()
    for _ in range(a.steps):
        st, nb = adv(dt, st, nb)
    torch.cuda.synchronize()
    el_prof = time.perf_counter() - t1
    ws.prof_enable(False)
    prof = ws.prof_read()
    overflow |= int(nb.error.code)
    # dominant kernel and its two roofline legs: algorithmic bytes against HBM
    # bandwidth, algorithmic flops against the float32 vector peak
    name, (tot_ms, cnt) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_s = tot_ms / cnt / 1e3
    pairs = nb.occupancy
    alg = dict(ALG_BYTES_3D)
    alg["nl_build"] = 12.0 + 4.0 * pairs / N
    per_launch = alg.get(name, float("nan")) * N
    achieved = per_launch / avg_s / 1e9
    flops = ALG_FLOPS_PER_PAIR_3D.get(name, float("nan")) * pairs
    tflops = flops / avg_s / 1e12
    # HBM bytes per launch of that kernel: rocprofv3 FETCH_SIZE / WRITE_SIZE of
    # this workload, scaled by a plain copy's counter readings (profiles/)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic_tgv3d_64cubed_fused_step.json")
    here = n_side == 64 and a.skin == "auto"
    if here and os.path.exists(tfile):
        for k, v in json.load(open(tfile))["kernels"].items():
            if k == f"k_{name}" or k.startswith(f"k_{name}<"):
                traffic = v["traffic_bytes"]
    # what bounds it, from the committed counters of this workload: the VALU
    # issue share against the HBM bandwidth share of the measured traffic
    counters = None
    if here and os.path.exists(COUNTERS):
        for k, v in json.load(open(COUNTERS))["kernels"].items():
            if k == f"k_{name}" or k.startswith(f"k_{name}<"):
                counters = v
    hbm_leg = {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
               "alg_bytes_per_launch": per_launch,
               "traffic_frac_of_peak": (traffic / avg_s / 1e9 / HBM_PEAK_GBS
                                        if traffic else None)}
    valu_leg = {"achieved": tflops, "peak": FP32_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": tflops / FP32_VALU_TFLOPS, "alg_flops_per_launch": flops,
                "issue_frac": counters["valu_issue_share"] if counters else None}
    bound = "hbm"
    if counters and counters["valu_issue_share"] > (hbm_leg["traffic_frac_of_peak"] or 0.0):
        bound = "valu"
    main_leg = valu_leg if bound == "valu" else hbm_leg
    roofline = {"bound": bound, "kernel": name, "achieved": main_leg["achieved"],
                "peak": main_leg["peak"], "unit": main_leg["unit"], "frac": main_leg["frac"],
                "traffic": traffic, "us_mean": avg_s * 1e6,
                "timing": "HIP events around every library launch, in a second pass of as "
                          "many steps (the events add host gaps between kernels: see "
                          "ms_per_step_profiled_pass; kernel durations agree with the "
                          "rocprofv3 trace in profiles/)",
                "hbm": hbm_leg, "valu": valu_leg,
                "counters": ({"source": os.path.relpath(COUNTERS, ROOT),
                              "valu_issue_share": counters["valu_issue_share"],
                              "ta_busy_share": counters.get("ta_busy_share")}
                             if counters else None)}
    kernels = {k: {"ms_total": round(v[0], 4), "launches": v[1],
                   "us_mean": round(v[0] / v[1] * 1e3, 2)} for k, v in
               sorted(prof.items(), key=lambda kv: -kv[1][0])}
    out = {
        "metric": "particle-steps/s fwd & fwd+adjoint, 3D TGV 262k particles; % of HBM peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 3D TGV lattice initial condition (sin/cos velocity field)",
        "config": {"workload": f"3D Taylor-Green vortex, {n_side}^3 = {N} particles, tvf=1, "
                               f"QSK h=dx, Re=1600, c0=10, dt={dt:.6g}, fused si_euler step",
                   "particles": N, "pairs_per_step": pairs, "dt": dt,
                   "verlet_skin": ws.skin, "xcd_remap": a.xcd_remap},
        "roofline": roofline,
        "kernels": kernels,
        "ms_per_step_profiled_pass": el_prof / a.steps * 1e3,
        "partition_error_code": overflow,
        "fwd_adjoint": None,
        "fwd_adjoint_3d": None,
    }
    if not a.no_adjoint:
        out["fwd_adjoint"] = fwd_adjoint_leg()
        out["fwd_adjoint_3d"] = fwd_adjoint_3d_leg(n_side)
    out["ensemble"] = None if a.no_ensemble else ensemble_leg(steps=a.ensemble_steps)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_sample_side)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

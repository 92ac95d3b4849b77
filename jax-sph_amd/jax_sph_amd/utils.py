"""Particle tags, lattice generators and statistics (``jax_sph/utils.py`` subset).

The lattice/wall generators run on the host once per case and reproduce the
reference's float32 rounding: positions are ``fl32(fl32(i + 0.5) * fl32(dx))``
(``utils.py:35-53``) and wall blocks are shifted by ``fl32(offset)``.
"""
from __future__ import annotations

import enum

import numpy as np
import torch


class Tag(enum.IntEnum):
    """utils.py:22-29."""

    PAD_VALUE = -1
    FLUID = 0
    SOLID_WALL = 1
    MOVING_WALL = 2
    DIRICHLET_WALL = 3


wall_tags = np.array([t.value for t in Tag if "WALL" in t.name])  # utils.py:32


def _dt(dtype):
    return np.dtype(dtype).type


def pos_init_cartesian_2d(box_size, dx: float, dtype=np.float32):
    """utils.py:35-45: cell-centred lattice, k = iy * nx + ix."""
    t = _dt(dtype)
    n = np.array((np.asarray(box_size) / dx).round(), dtype=int)
    grid = np.meshgrid(range(n[0]), range(n[1]), indexing="xy")
    idx = np.vstack(list(map(np.ravel, grid))).T
    return ((idx + t(0.5)).astype(dtype) * t(dx)).astype(dtype)


def pos_init_cartesian_3d(box_size, dx: float, dtype=np.float32):
    """utils.py:48-53: k = (iy * nx + ix) * nz + iz."""
    t = _dt(dtype)
    n = np.array((np.asarray(box_size) / dx).round(), dtype=int)
    grid = np.meshgrid(range(n[0]), range(n[1]), range(n[2]), indexing="xy")
    idx = np.vstack(list(map(np.ravel, grid))).T
    return ((idx + t(0.5)).astype(dtype) * t(dx)).astype(dtype)


def _shift(r, off, dtype):
    return (r + np.asarray(off, dtype=np.float64).astype(dtype)).astype(dtype)


def pos_box_2d(fluid_box, dx: float, n_walls: int = 3, dtype=np.float32):
    """utils.py:56-77: walls left, bottom, right, top."""
    dxn = n_walls * dx
    vertical = pos_init_cartesian_2d(np.array([dxn, fluid_box[1] + 2 * dxn]), dx, dtype)
    horiz = pos_init_cartesian_2d(np.array([fluid_box[0], dxn]), dx, dtype)
    wall_l = vertical.copy()
    wall_b = _shift(horiz, [dxn, 0.0], dtype)
    wall_r = _shift(vertical, [fluid_box[0] + dxn, 0.0], dtype)
    wall_t = _shift(horiz, [dxn, fluid_box[1] + dxn], dtype)
    return np.concatenate([wall_l, wall_b, wall_r, wall_t])


def pos_box_3d(fluid_box, dx: float, n_walls: int = 3, z_periodic: bool = True,
               dtype=np.float32):
    """utils.py:80-118."""
    dxn = n_walls * dx
    vertical = pos_init_cartesian_3d(np.array([dxn, fluid_box[1] + 2 * dxn, fluid_box[2]]), dx,
                                     dtype)
    horiz = pos_init_cartesian_3d(np.array([fluid_box[0], dxn, fluid_box[2]]), dx, dtype)
    res = np.concatenate([vertical.copy(), _shift(horiz, [dxn, 0.0, 0.0], dtype),
                          _shift(vertical, [fluid_box[0] + dxn, 0.0, 0.0], dtype),
                          _shift(horiz, [dxn, fluid_box[1] + dxn, 0.0], dtype)])
    if not z_periodic:
        res = _shift(res, [0.0, 0.0, dxn], dtype)
        front = pos_init_cartesian_3d(
            np.array([fluid_box[0] + 2 * dxn, fluid_box[1] + 2 * dxn, dxn]), dx, dtype)
        res = np.concatenate([res, front, _shift(front, [0.0, 0.0, fluid_box[2] + dxn], dtype)])
    return res


def get_noise_masked(shape, mask, generator: torch.Generator, std: float):
    """utils.py:121-125 with a torch generator instead of a JAX PRNG key."""
    noise = std * torch.randn(shape, generator=generator, dtype=torch.float32)
    m = torch.as_tensor(np.asarray(mask))[:, None]
    return torch.where(m, noise, torch.zeros_like(noise))


def get_ekin(state, dx: float) -> float:
    """utils.py:128-133: 0.5 * sum(v_fluid^2) * dx^d from ``state["v"]``."""
    v = state["v"]
    fluid = (state["tag"] == Tag.FLUID)[:, None]
    vw = torch.where(fluid, v, torch.zeros_like(v)).double()
    return 0.5 * float(torch.square(vw).sum()) * dx ** v.shape[1]


def get_array_stats(state, var: str = "u", operation: str = "max"):
    """utils.py:136-153 (Euclidean norm for vectors)."""
    x = state[var]
    if x.dim() > 1:
        x = torch.sqrt(torch.square(x).sum(dim=1))
    return float({"min": torch.min, "max": torch.max, "mean": torch.mean}[operation](x))


def get_stats(state, props, dx: float):
    """utils.py:156-166."""
    res = {}
    for prop in props:
        if prop == "Ekin":
            res[prop] = get_ekin(state, dx)
        else:
            var, op = prop.split("_")
            res[prop] = get_array_stats(state, var, op)
    return res


class Logger:
    """utils.py:278-296."""

    def __init__(self, dt, dx, print_props, sequence_length) -> None:
        self.dt, self.dx = dt, dx
        self.print_props = print_props
        self.sequence_length = sequence_length
        self.digits = len(str(sequence_length))

    def print_stats(self, state, step):
        t_ = (step + 1) * self.dt
        stats = get_stats(state, self.print_props, self.dx)
        s = ", ".join([f"{k}={v:.5f}" for k, v in stats.items()])
        print(f"{str(step).zfill(self.digits)}/{self.sequence_length}, t={t_:.4f}, {s}")


def sph_interpolator(cfg, src_state, prop_type: str = "vector"):
    """utils.py:299-443: kernel interpolation of a state onto arbitrary points.

    The reference reads the source state from an h5 file (``src_path``); this
    package has no h5 reader, so the source state dict is passed instead, and
    the returned function takes a state dict in place of a path:
    ``interp(state, r_target, prop="u", dim_ind=0)`` for vectors,
    ``interp(state, r_target, prop="p")`` for scalars.

    As in the reference, the neighbour list (pairs within 3 dx) is built once
    from the source positions; wall particles first take the generalized wall
    BC values over their fluid neighbours (utils.py:316-337: 2 x - x_wall for
    vectors, x_wall for scalars), then every target point gets the
    Shepard-normalised Quintic sum over all particles (plain Euclidean
    distance, utils.py:392-401).  Runs on the state's device in float32.
    """
    from . import partition
    from .jax_md import space
    from .kernel import QuinticKernel

    if prop_type not in ("vector", "scalar"):
        raise ValueError(f"prop_type must be 'vector' or 'scalar', not {prop_type!r}")
    r0 = src_state["r"]
    dev = r0.device
    N, dim = r0.shape
    dx = float(cfg.case.dx)
    eps = np.float32(np.finfo(float).eps)
    kernel_fn = QuinticKernel(h=dx, dim=dim)
    box_size = np.array(cfg.case.bounds)[:, 1]
    if np.array(cfg.case.pbc).sum() == 0:
        raise NotImplementedError("free-space interpolation is not provided")
    displacement_fn, _ = space.periodic(side=box_size)
    neighbor_fn = partition.neighbor_list(
        displacement_fn, box_size, r_cutoff=3 * dx, dr_threshold=3 * dx * 0.25,
        capacity_multiplier=1.25, mask_self=False, format=partition.Sparse)
    idx = neighbor_fn.allocate(r0).idx.long()
    keep = idx[0] < N
    i_s, j_s = idx[0][keep], idx[1][keep]
    wall = torch.as_tensor(wall_tags, device=dev)

    def _wall_weights(state):
        r = state["r"].float()
        dist = space.distance(displacement_fn(r[i_s], r[j_s]))
        w = kernel_fn.w(dist)
        w_fluid = w * (state["tag"][j_s] == Tag.FLUID).float()
        w_sum = torch.zeros(N, device=dev).index_add_(0, i_s, w_fluid)
        return w_fluid, w_sum, torch.isin(state["tag"], wall)

    def _interpolate(state, x, r_target):
        r = state["r"].float()
        rt = torch.as_tensor(np.asarray(r_target), dtype=torch.float32, device=dev)
        out = []
        for a in range(0, rt.shape[0], 256):  # bounded (targets x N) blocks
            d = torch.sqrt(((rt[a:a + 256, None, :] - r[None, :, :]) ** 2).sum(-1))
            w = kernel_fn.w(d)
            w_norm = w.sum(dim=1) * dx ** dim
            out.append((w * x[None, :]).sum(dim=1) * dx ** dim / w_norm)
        return torch.cat(out)

    def interp_vel(state, r_target, prop: str = "u", dim_ind: int = 0):
        w_fluid, w_sum, mask_bc = _wall_weights(state)
        vel = state[prop].float()
        x_wall = torch.zeros_like(vel).index_add_(0, i_s, w_fluid[:, None] * vel[j_s])
        x_wall = x_wall / (w_sum[:, None] + eps)
        vel = torch.where(mask_bc[:, None], 2 * vel - x_wall, vel)
        return _interpolate(state, vel[:, dim_ind], r_target)

    def interp_scalar(state, r_target, prop: str = "p"):
        w_fluid, w_sum, mask_bc = _wall_weights(state)
        x = state[prop].float()
        x_wall = torch.zeros_like(x).index_add_(0, i_s, w_fluid * x[j_s]) / (w_sum + eps)
        x = torch.where(mask_bc, x_wall, x)
        return _interpolate(state, x, r_target)

    return interp_vel if prop_type == "vector" else interp_scalar

"""Oracle: periodic space, cell list and sparse neighbour list (numpy restatement).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, never by the product package.

Restates, statement by statement, the reference's

* ``jax_sph/jax_md/space.py:170-209``   periodic displacement / shift / distance
* ``jax_sph/jax_md/partition.py:114-428`` cell list (hash, stable sort, slot trick,
  occupancy, ``CELL_LIST_OVERFLOW``)
* ``jax_sph/jax_md/partition.py:723-1026`` ``neighbor_list`` (vmap backend: 3^d
  rolled candidate cells, fp32 prune, cumsum compaction, capacity rules,
  ``NEIGHBOR_LIST_OVERFLOW``, rebuild condition)
* ``jax_sph/partition.py:55-353`` ``_scan_neighbor_list`` (chunked variant:
  per-chunk edge budget, ``CELL_LIST_OVERFLOW`` on chunk overflow, final argsort)

in numpy with an explicit dtype (float32 is the reference's ``dtype=float32``
path; float64 mirrors ``jax_enable_x64``).  The idx layout (row 0 receivers,
row 1 senders, sender-sorted, candidate order inside a sender, padding N) is
reproduced exactly, including the ``(k mod capacity)`` slot rotation inside a
cell (``jax_md/partition.py:386-391``).

How each part is pinned (tests/test_oracle_nl.py):
* pair sets: the reference's golden targets (``tests/test_neighbors.py:87-121``)
  and an O(N^2) brute force over all periodic images on random inputs;
* capacity: the closed form ``int(fl32(occ*1.25) + N*extra)`` on hand-counted
  lattices, the overflow bits on forced-overflow inputs;
* geometry: minimum-image identities (antisymmetry, |dR| <= side/2).

Known divergence kept on purpose by the product and flagged here: a particle
whose ``int32(r/cell)`` equals the number of cells (r == side after rounding) is
hashed into a wrong or out-of-range cell by the reference; ``clamp=True``
clamps it into the last cell, which is what the product does.  With
``clamp=False`` (reference behaviour) the oracle raises ``EdgeOfBoxError`` so a
test never silently compares such a case.
"""
from __future__ import annotations

import dataclasses
import itertools
from typing import Optional

import numpy as np

NEIGHBOR_LIST_OVERFLOW = 1
CELL_LIST_OVERFLOW = 2
CELL_SIZE_TOO_SMALL = 4
MALFORMED_BOX = 8


class EdgeOfBoxError(RuntimeError):
    """A particle hashes to cell index == cells_per_side (reference drops/aliases it)."""


# --------------------------------------------------------------------------- space
def jmod(x, y):
    """``jnp.mod``: C fmod plus the sign fix towards the divisor."""
    r = np.fmod(x, y)
    adj = (r != 0) & ((r < 0) != (y < 0))
    return np.where(adj, r + y, r)


def periodic_displacement(side, dR):
    """``jax_md/space.py:170-181``: mod(dR + side/2, side) - side/2."""
    half = side * side.dtype.type(0.5)
    return jmod(dR + half, side) - half


def periodic_shift(side, R, dR):
    """``jax_md/space.py:207-209``."""
    return jmod(R + dR, side)


def square_distance(dR):
    """``jax_md/space.py:184-192``; x^2 then y^2 then z^2, no fused multiply-add."""
    acc = dR[..., 0] * dR[..., 0]
    for k in range(1, dR.shape[-1]):
        acc = acc + dR[..., k] * dR[..., k]
    return acc


def distance(dR):
    """``jax_md/space.py:195-204`` with ``safe_mask`` (0 at 0)."""
    d2 = square_distance(dR)
    return np.where(d2 > 0, np.sqrt(np.where(d2 > 0, d2, 0)), 0).astype(dR.dtype)


# ------------------------------------------------------------------- cell geometry
def cell_dimensions(box, minimum_cell_size):
    """``jax_md/partition.py:114-156`` for an array box (float arithmetic in box dtype)."""
    box = np.asarray(box)
    cps = np.floor(box / minimum_cell_size)
    cell_size = box / cps
    cps_i = cps.astype(np.int32)
    if np.any(cps_i < 3):
        raise ValueError(
            "Box must be at least 3x the size of the grid spacing in each dimension."
        )
    return cell_size, cps_i, int(np.prod(cps_i.astype(np.int64)))


def hash_multipliers(cps):
    """``jax_md/partition.py:177-187``: [1, nx, nx*ny]."""
    out = np.ones(len(cps), np.int64)
    for d in range(1, len(cps)):
        out[d] = out[d - 1] * int(cps[d - 1])
    return out


def particle_cell_index(position, cell_size, cps, clamp):
    """``int32(position / cell_size)`` per dimension (partition.py:367)."""
    q = (position / cell_size.astype(position.dtype)).astype(np.int32)
    if clamp:
        q = np.minimum(np.maximum(q, 0), cps.astype(np.int32) - 1)
    elif np.any(q >= cps) or np.any(q < 0):
        raise EdgeOfBoxError("particle hashed outside [0, cells_per_side)")
    return q


def count_cell_filling(position, box, minimum_cell_size, clamp=False):
    """``jax_md/partition.py:159-174`` (uses its own grid derived from the cell size)."""
    cell_size, cps, cell_count = cell_dimensions(box, minimum_cell_size)
    idx = particle_cell_index(position, cell_size, cps, clamp)
    h = (idx.astype(np.int64) * hash_multipliers(cps)).sum(axis=1)
    return np.bincount(h, minlength=cell_count)[:cell_count]


# ----------------------------------------------------------------------- cell list
@dataclasses.dataclass
class CellList:
    id_buffer: np.ndarray       # (cell_count, capacity) int32, empty = N
    particle_cell: np.ndarray   # (N,) flat cell hash of every particle
    occupancy: np.ndarray       # (cell_count,) particles per cell
    overflow: bool
    capacity: int
    cell_size: np.ndarray
    cps: np.ndarray


def cell_list_build(position, box, minimum_cell_size, capacity=None,
                    buffer_size_multiplier=1.25, extra_capacity=0, clamp=False):
    """``jax_md/partition.py:294-413`` (allocate when ``capacity`` is None)."""
    N = position.shape[0]
    cell_size, cps, cell_count = cell_dimensions(box, minimum_cell_size)
    if capacity is None:
        # _estimate_cell_capacity (partition.py:195-199) re-derives a grid from
        # the already rounded cell size.
        filling = count_cell_filling(position, box, cell_size, clamp=clamp)
        capacity = int(int(filling.max()) * buffer_size_multiplier) + extra_capacity
    idx = particle_cell_index(position, cell_size, cps, clamp)
    hashes = (idx.astype(np.int64) * hash_multipliers(cps)).sum(axis=1)
    sort_map = np.argsort(hashes, kind="stable")
    sorted_hash = hashes[sort_map]
    slot = sorted_hash * capacity + (np.arange(N) % capacity)
    flat = np.full(cell_count * capacity, N, np.int32)
    flat[slot] = sort_map.astype(np.int32)  # later duplicates win, like XLA CPU
    occupancy = np.bincount(hashes, minlength=cell_count)[:cell_count]
    overflow = bool(occupancy.max() > capacity) if N else False
    return CellList(flat.reshape(cell_count, capacity), hashes, occupancy, overflow,
                    capacity, cell_size, cps)


def neighbor_cell_offsets(dim):
    """``_neighboring_cells`` order (partition.py:190-192), own cell first.

    The rolled copies of partition.py:202-224 act on the buffer reshaped to
    (nz, ny, nx, cap): the first offset component rolls z (y in 2D).  A shift of
    -1 brings in the cell at +1, so the candidate cell is ``cell - offset``.
    Returns offsets as (dx, dy[, dz]) of the candidate cell relative to the own.
    """
    out = [np.zeros(dim, np.int64)]
    for dindex in itertools.product((-1, 0, 1), repeat=dim):
        d = np.array(dindex, np.int64)
        if np.all(d == 0):
            continue
        out.append(-d[::-1])  # (x, y[, z]) order, candidate = own - shift
    return out


def cell_candidates(cl: CellList, N):
    """``cell_list_candidate_fn`` (partition.py:832-856): (N, 3^d*cap) ids."""
    cps = cl.cps.astype(np.int64)
    dim = len(cps)
    cell_count = cl.id_buffer.shape[0]
    mult = hash_multipliers(cps)
    c = np.arange(cell_count, dtype=np.int64)
    coord = np.stack([(c // mult[d]) % cps[d] for d in range(dim)], axis=1)
    blocks = []
    for off in neighbor_cell_offsets(dim):
        nc = ((coord + off) % cps * mult).sum(axis=1)
        blocks.append(cl.id_buffer[nc])
    cell_idx = np.concatenate(blocks, axis=1)  # (cell_count, 3^d*cap)
    out = np.zeros((N + 1, cell_idx.shape[1]), np.int32)
    flat_ids = cl.id_buffer.reshape(-1)
    rows = np.repeat(np.arange(cell_count), cl.capacity)
    out[flat_ids] = cell_idx[rows]
    return out[:-1]


# -------------------------------------------------------------------- neighbour list
@dataclasses.dataclass
class NeighborList:
    idx: np.ndarray
    reference_position: np.ndarray
    error: int
    cell_list_capacity: Optional[int]
    max_occupancy: int
    occupancy: int

    @property
    def did_buffer_overflow(self):
        return self.error & (NEIGHBOR_LIST_OVERFLOW | CELL_LIST_OVERFLOW)


class NeighborSearch:
    """``jax_md.partition.neighbor_list`` (vmap) and ``_scan_neighbor_list`` (scan).

    Sparse / OrderedSparse formats only (Dense is not on the SPH path).
    """

    def __init__(self, box, r_cutoff, dr_threshold=0.0, capacity_multiplier=1.25,
                 disable_cell_list=False, mask_self=True, ordered=False,
                 backend="jaxmd_vmap", num_partitions=1, dtype=np.float32,
                 clamp=False):
        t = np.dtype(dtype).type
        self.dtype = np.dtype(dtype)
        self.box = np.asarray(box, dtype=self.dtype).reshape(-1)
        self.cutoff = t(t(r_cutoff) + t(dr_threshold))          # partition.py:815-820
        self.cutoff_sq = t(self.cutoff * self.cutoff)
        self.threshold_sq = t(t(t(dr_threshold) / t(2)) ** 2)
        self.capacity_multiplier = capacity_multiplier
        self.disable_cell_list = disable_cell_list
        self.mask_self = mask_self
        self.ordered = ordered
        self.backend = backend
        self.num_partitions = num_partitions
        self.clamp = clamp
        if backend == "jaxmd_scan":
            assert not disable_cell_list and not ordered
            assert np.all(self.cutoff < self.box / 3.0)

    # metric used by prune and by the rebuild test
    def metric_sq(self, Ra, Rb):
        return square_distance(periodic_displacement(self.box, Ra - Rb))

    def _use_cells(self):
        return (not self.disable_cell_list) and bool(np.all(self.cutoff < self.box / 3.0))

    def _candidates(self, position, cl_capacity, extra_capacity):
        N = position.shape[0]
        if self._use_cells():
            cl = cell_list_build(position, self.box, self.cutoff, cl_capacity,
                                 self.capacity_multiplier, extra_capacity, self.clamp)
            return cell_candidates(cl, N), cl
        return np.broadcast_to(np.arange(N, dtype=np.int32)[None], (N, N)).copy(), None

    def _prune(self, position, cand, sender_offset=0):
        N = position.shape[0]
        rows = cand.shape[0]
        sender = np.broadcast_to(
            (np.arange(rows, dtype=np.int64) + sender_offset)[:, None], cand.shape
        ).reshape(-1)
        recv = cand.reshape(-1).astype(np.int64)
        valid = recv < N
        rs = np.minimum(recv, N - 1)
        ss = np.minimum(sender, N - 1)
        d2 = self.metric_sq(position[ss], position[rs])
        mask = (d2 < self.cutoff_sq) & valid
        if self.ordered:
            mask &= recv < sender
        keep = np.nonzero(mask)[0]
        return recv[keep].astype(np.int32), sender[keep].astype(np.int32), len(keep)

    def _capacity(self, occupancy, N, extra_capacity, width):
        t = self.dtype.type
        if self.dtype == np.float32:
            # int(occ * mult + extra): int32 * weak float -> float32, + weak int -> float32
            m = int(t(t(t(occupancy) * t(self.capacity_multiplier)) + t(N * extra_capacity)))
        else:
            m = int(occupancy * self.capacity_multiplier + N * extra_capacity)
        m = min(m, width)
        if self.ordered:
            limit = N * (N - 1) // 2
        else:
            limit = N * (N - 1) if self.mask_self else N * N
        return min(m, limit)

    def _build(self, position, cl_capacity=None, max_occupancy=None, extra_capacity=0,
               error=0):
        position = np.asarray(position, dtype=self.dtype)
        N = position.shape[0]
        cand, cl = self._candidates(position, cl_capacity, extra_capacity)
        if cl is not None and cl.overflow:
            error |= CELL_LIST_OVERFLOW
        if self.mask_self:
            cand = np.where(cand == np.arange(N, dtype=np.int32)[:, None], N, cand)
        if self.backend == "jaxmd_scan":
            return self._build_scan(position, cand, cl, max_occupancy, extra_capacity,
                                    error)
        recv, send, occ = self._prune(position, cand)
        width = cand.size
        if max_occupancy is None:
            max_occupancy = self._capacity(occ, N, extra_capacity, width)
        idx = np.full((2, max_occupancy), N, np.int32)
        k = min(occ, max_occupancy)
        idx[0, :k] = recv[:k]
        idx[1, :k] = send[:k]
        if occ > max_occupancy:
            error |= NEIGHBOR_LIST_OVERFLOW
        return NeighborList(idx, position, error,
                            None if cl is None else cl.capacity, max_occupancy, occ)

    def _build_scan(self, position, cand, cl, max_occupancy, extra_capacity, error):
        # jax_sph/partition.py:208-311
        N, dim = position.shape
        considered = cand.shape[1]
        n_sub = int(np.ceil(N / self.num_partitions))
        vol = 0.34907 if dim == 2 else 0.15514
        budget = int(n_sub * considered * vol * self.capacity_multiplier)
        rows_r, rows_s = [], []
        occ = 0
        for p in range(self.num_partitions):
            lo, hi = p * n_sub, min((p + 1) * n_sub, N)
            r, s, c = self._prune(position, cand[lo:hi], sender_offset=lo)
            occ += c
            if c > budget:
                error |= CELL_LIST_OVERFLOW
            rr = np.full(budget, N, np.int32)
            ss = np.full(budget, N, np.int32)
            k = min(c, budget)
            rr[:k], ss[:k] = r[:k], s[:k]
            rows_r.append(rr)
            rows_s.append(ss)
        recv = np.concatenate(rows_r)
        send = np.concatenate(rows_s)
        order = np.argsort(send, kind="stable")
        recv, send = recv[order], send[order]
        if max_occupancy is None:
            max_occupancy = self._capacity(occ, N, extra_capacity, recv.shape[0])
        idx = np.stack([recv[:max_occupancy], send[:max_occupancy]])
        if idx.shape[1] < max_occupancy:  # cannot happen: capped by width
            pad = np.full((2, max_occupancy - idx.shape[1]), N, np.int32)
            idx = np.concatenate([idx, pad], axis=1)
        if occ > max_occupancy:
            error |= NEIGHBOR_LIST_OVERFLOW
        return NeighborList(idx, position, error, cl.capacity, max_occupancy, occ)

    # public API mirroring NeighborListFns
    def allocate(self, position, extra_capacity=0):
        return self._build(position, extra_capacity=extra_capacity)

    def update(self, position, nbrs: NeighborList):
        position = np.asarray(position, dtype=self.dtype)
        d = self.metric_sq(position, nbrs.reference_position)
        if not np.any(d > self.threshold_sq):                    # partition.py:1012-1018
            return nbrs
        return self._build(position, nbrs.cell_list_capacity, nbrs.max_occupancy,
                           error=nbrs.error)


# --------------------------------------------------------------------- brute force
def brute_force_pairs(position, box, cutoff, mask_self=True, dtype=np.float32):
    """All (receiver, sender) pairs with fp32 min-image d^2 < fl(cutoff)^2, O(N^2).

    Independent of the cell list: it is the set the cell-list search must
    reproduce whenever every pair within the cutoff lies in neighbouring cells.
    """
    t = np.dtype(dtype).type
    position = np.asarray(position, dtype)
    box = np.asarray(box, dtype)
    c = t(cutoff)
    c2 = t(c * c)
    N = position.shape[0]
    dR = periodic_displacement(box, position[:, None, :] - position[None, :, :])
    d2 = square_distance(dR)  # d2[s, r] = |r_s - r_r|^2
    m = d2 < c2
    if mask_self:
        np.fill_diagonal(m, False)
    s, r = np.nonzero(m)
    return np.stack([r, s]).astype(np.int32)


def sorted_pairs(idx, N):
    """Real pairs of an idx array as a lexicographically (sender, receiver) sorted array."""
    idx = np.asarray(idx)
    keep = idx[0] < N
    real = idx[:, keep].astype(np.int64)
    order = np.lexsort((real[0], real[1]))
    return real[:, order]

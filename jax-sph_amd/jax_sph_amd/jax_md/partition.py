"""Neighbour lists on the HIP library (drop-in for ``jax_sph/jax_md/partition.py``).

Same call signatures and semantics as the reference's ``neighbor_list``
(``jax_md/partition.py:723-1026``) for the Sparse and OrderedSparse formats:

* ``allocate`` sizes everything on the host (eager, like the reference):
  the emulated cell capacity ``int(max_fill * capacity_multiplier) + extra``
  (partition.py:195-199, 316-319) and ``max_occupancy = int(occ *
  capacity_multiplier + N * extra)`` capped like partition.py:956-970;
* ``update`` never resizes; it sets ``NEIGHBOR_LIST_OVERFLOW`` when the pair
  count exceeds ``max_occupancy`` and ``CELL_LIST_OVERFLOW`` when a cell holds
  more particles than the emulated capacity; error bits are sticky;
* ``NeighborList.idx`` is int32 ``[2, max_occupancy]`` (receivers, senders),
  sender-sorted and padded with N.  It is written by a HIP kernel the first
  time it is read; the step itself works on the per-particle pair buffers.

Differences kept on purpose: inside one sender the receivers are listed in
cell-strip order instead of the reference's rolled-copy order (pair *sets* are
bit-exact); a particle exactly on the upper box face is binned into the last
cell instead of being aliased into a wrong cell.  Each ``update`` writes into
the other of two buffer banks, so the list an update started from stays
readable; a list two updates old raises ``StaleNeighborListError``.

Skin lists (extra keyword ``skin``, default ``"auto"`` = 0.08 * cutoff; 0
turns them off): when the fused ``si_euler`` step advances a list, it keeps
pairs within ``cutoff + skin`` and rebuilds them on the device only after a
particle has moved more than ``0.49 * skin``; the reference predicate is
applied to every pair on every step, so pair sets, occupancy and error bits
are those of a per-step rebuild.  A list produced that way is stale after the
next fused step.  If a skin list outgrows its buffer the step reports
``NEIGHBOR_LIST_OVERFLOW`` and the re-allocation sizes the buffer from the
observed count.
"""
from __future__ import annotations

import ctypes
import enum
from typing import Callable, Optional

import numpy as np
import torch

from .. import _lib
from .space import PeriodicDisplacement

# Verlet skin of the fused step's lists as a multiple of the cutoff (skin
# "auto").  Swept on the 3D Taylor-Green vortex at 64^3: 0.07, 0.08, 0.09, 0.1,
# 0.11 x cutoff give 0.293, 0.285, 0.289, 0.290, 0.291 ms per step over 100
# steps and 0.304, 0.294, 0.296, 0.301, 0.298 ms over 500: wider skins list
# more pairs, narrower ones rebuild more often.
SKIN_FACTOR = 0.08

PEC_NONE = 0


class PartitionErrorCode(enum.IntEnum):
    """partition.py:434-457."""

    NONE = 0
    NEIGHBOR_LIST_OVERFLOW = 1 << 0
    CELL_LIST_OVERFLOW = 1 << 1
    CELL_SIZE_TOO_SMALL = 1 << 2
    MALFORMED_BOX = 1 << 3


PEC = PartitionErrorCode


class NeighborListFormat(enum.Enum):
    """partition.py:579-599."""

    Dense = 0
    Sparse = 1
    OrderedSparse = 2


Dense = NeighborListFormat.Dense
Sparse = NeighborListFormat.Sparse
OrderedSparse = NeighborListFormat.OrderedSparse


def is_sparse(fmt) -> bool:
    return fmt is Sparse or fmt is OrderedSparse


class StaleNeighborListError(RuntimeError):
    """The buffers behind this NeighborList were reused by a later update."""


class PartitionError:
    """partition.py:463-499; ``code`` is a 0-d uint8 device tensor."""

    def __init__(self, code: torch.Tensor):
        self.code = code

    def update(self, bit, pred) -> "PartitionError":
        pred = torch.as_tensor(pred, device=self.code.device)
        return PartitionError(self.code | torch.where(pred, torch.tensor(int(bit), dtype=torch.uint8,
                                                                          device=self.code.device),
                                                      torch.zeros((), dtype=torch.uint8,
                                                                  device=self.code.device)))

    def __str__(self) -> str:
        c = int(self.code)
        if not c:
            return ""
        if c & PEC.NEIGHBOR_LIST_OVERFLOW:
            return "Partition Error: Neighbor list buffer overflow."
        if c & PEC.CELL_LIST_OVERFLOW:
            return "Partition Error: Cell list buffer overflow"
        if c & PEC.CELL_SIZE_TOO_SMALL:
            return "Partition Error: Cell size too small"
        if c & PEC.MALFORMED_BOX:
            return "Partition Error: Incorrect box format. Expecting upper triangular."
        raise ValueError(f"Unexpected Error Code {c}.")

    __repr__ = __str__


class NeighborList:
    """partition.py:620-671.  Dynamic state lives in device buffers of a workspace bank."""

    def __init__(self, search: "_NeighborSearch", ws: "_lib.Workspace", bank: int,
                 info: torch.Tensor, reference_position: torch.Tensor,
                 cell_list_capacity: Optional[int], max_occupancy: int, update_fn: Callable):
        self._search = search
        self._ws = ws
        self._bank = bank
        self._gen = ws.generation[bank]
        self._info = info
        self._idx = None
        self.reference_position = reference_position
        self.cell_list_capacity = cell_list_capacity
        self.max_occupancy = int(max_occupancy)
        self.format = search.format
        self.cell_size = search.cutoff if search.use_cells else None
        self.cell_list_fn = search if search.use_cells else None
        self.update_fn = update_fn

    # --- dynamic fields
    def _check_live(self):
        if self._ws.generation[self._bank] != self._gen:
            raise StaleNeighborListError(
                "this NeighborList's buffers were reused by a later update; read .idx "
                "before updating twice")

    @property
    def idx(self) -> torch.Tensor:
        if self._idx is None:
            self._check_live()
            self._idx = self._search.write_idx(self)
        return self._idx

    @property
    def error(self) -> PartitionError:
        return PartitionError(self._info[0].to(torch.uint8))

    @property
    def did_buffer_overflow(self) -> torch.Tensor:
        return self.error.code & (PEC.NEIGHBOR_LIST_OVERFLOW | PEC.CELL_LIST_OVERFLOW)

    @property
    def cell_size_too_small(self) -> torch.Tensor:
        return self.error.code & PEC.CELL_SIZE_TOO_SMALL

    @property
    def malformed_box(self) -> torch.Tensor:
        return self.error.code & PEC.MALFORMED_BOX

    # --- extras of this implementation
    @property
    def occupancy(self) -> int:
        """Number of real pairs found by the last build (host sync)."""
        return int(self._info[1])

    @property
    def info(self) -> torch.Tensor:
        return self._info

    @property
    def bank(self) -> int:
        return self._bank

    @property
    def workspace(self) -> "_lib.Workspace":
        return self._ws

    def update(self, position, **kwargs) -> "NeighborList":
        return self.update_fn(position, self, **kwargs)


class NeighborListFns:
    """partition.py:674-717."""

    def __init__(self, allocate: Callable, update: Callable):
        self.allocate = allocate
        self.update = update

    def __call__(self, position, neighbors=None, extra_capacity: int = 0, **kwargs):
        if neighbors is None:
            return self.allocate(position, extra_capacity, **kwargs)
        return self.update(position, neighbors, **kwargs)

    def __iter__(self):
        return iter((self.allocate, self.update))


def _f32(x):
    return np.float32(x)


class _NeighborSearch:
    """Host side of one neighbour-list configuration; owns the HIP workspace."""

    def __init__(self, displacement_or_metric, box, r_cutoff, dr_threshold=0.0,
                 capacity_multiplier=1.25, disable_cell_list=False, mask_self=True,
                 custom_mask_function=None, fractional_coordinates=False, format=Sparse,
                 backend="jaxmd_vmap", num_partitions=1, skin="auto"):
        if not isinstance(format, NeighborListFormat):
            raise ValueError(f"Neighbor list format must be a member of NeighborListFormat "
                             f"found {format}.")
        if format is Dense:
            raise NotImplementedError("Dense neighbour lists are not on the SPH step; use Sparse")
        if custom_mask_function is not None:
            raise NotImplementedError("custom_mask_function is not supported")
        if fractional_coordinates:
            raise NotImplementedError("fractional coordinates are not supported")
        if displacement_or_metric is not None and not isinstance(displacement_or_metric,
                                                                 PeriodicDisplacement):
            raise NotImplementedError("only space.periodic displacements are supported "
                                      "(case_setup.py:133 always uses them)")
        self.box_arg = box
        self.r_cutoff = float(r_cutoff)
        self.dr_threshold = float(dr_threshold)
        self.cutoff = _f32(_f32(r_cutoff) + _f32(dr_threshold))          # partition.py:815-821
        self.cutoff_sq = _f32(self.cutoff * self.cutoff)
        self.threshold_sq = _f32(_f32(_f32(dr_threshold) / _f32(2)) ** 2)
        self.capacity_multiplier = float(capacity_multiplier)
        self.disable_cell_list = disable_cell_list
        self.mask_self = bool(mask_self)
        self.format = format
        self.backend = backend
        self.num_partitions = int(num_partitions)
        self.use_cells = None
        if skin == "auto":
            skin = SKIN_FACTOR * float(self.cutoff)
        self.skin = float(skin or 0.0)
        if self.skin < 0.0:
            raise ValueError("skin must be >= 0")
        self._skin_hint = 0
        self._last_step_info = None

    # ------------------------------------------------------------ geometry
    def _geometry(self, dim):
        box = np.asarray(self.box_arg, dtype=np.float32).reshape(-1)
        if box.size == 1:
            box = np.full(dim, box[0], np.float32)
        if box.size != dim:
            raise ValueError(f"box has {box.size} entries for {dim}-d positions")
        self.box = box
        self.use_cells = (not self.disable_cell_list) and bool(
            np.all(self.cutoff < box / np.float32(3.0)))              # partition.py:929
        if self.backend == "jaxmd_scan" and not self.use_cells:
            raise AssertionError("Don't use scan with very few cells")
        if self.use_cells:
            cps = np.floor(box / self.cutoff)                          # partition.py:128
            self.cell_size = (box / cps).astype(np.float32)
            self.cps = cps.astype(np.int32)
            if np.any(self.cps < 3):
                raise ValueError("Box must be at least 3x the size of the grid spacing in "
                                 "each dimension.")
            cps2 = np.floor(box / self.cell_size)                      # partition.py:159-174
            self.cell_size2 = (box / cps2).astype(np.float32)
            self.cps2 = cps2.astype(np.int32)
        else:
            self.cell_size = np.ones(dim, np.float32)
            self.cps = np.ones(dim, np.int32)

    def _config(self, n, dim, cap, width):
        c = _lib.NLConfig()
        c.dim, c.n = dim, n
        for d in range(3):
            c.box[d] = float(self.box[d]) if d < dim else 0.0
            c.cps[d] = int(self.cps[d]) if d < dim else 1
            c.cell_size[d] = float(self.cell_size[d]) if d < dim else 1.0
        c.cutoff, c.cutoff_sq = float(self.cutoff), float(self.cutoff_sq)
        c.use_cells = 1 if self.use_cells else 0
        c.cell_capacity = int(cap) if cap is not None else -1
        c.mask_self = 1 if self.mask_self else 0
        c.ordered = 1 if self.format is OrderedSparse else 0
        c.ell_width = int(width)
        return c

    # ------------------------------------------------------------ builds
    def allocate(self, position, extra_capacity: int = 0, ell_width: Optional[int] = None,
                 **kwargs) -> NeighborList:
        """partition.py:911-1021 (neighbor_list_fn with neighbors=None, then
        allocate_fn).  ``ell_width`` (not in the reference) caps the
        pair buffer at that many pairs per particle instead of the 3^d x cell
        capacity candidates; a particle with more pairs then raises
        NEIGHBOR_LIST_OVERFLOW (used by workspaces that hold many lists)."""
        dev = _lib.device()
        r = _lib.as_f32(position, dev)
        n, dim = r.shape
        if dim not in (2, 3):
            raise ValueError(f"Cell list spatial dimension must be 2 or 3. Found {dim}.")
        self._geometry(dim)
        L = _lib.lib()
        st = _lib.stream()
        info = torch.zeros(8, dtype=torch.int64, device=dev)
        cap = None
        if self.use_cells:
            # capacity estimate on the grid re-derived from the rounded cell size
            counts = torch.empty(int(np.prod(self.cps2.astype(np.int64))), dtype=torch.int32,
                                 device=dev)
            cs2 = (ctypes.c_float * 3)(*([float(x) for x in self.cell_size2] + [1.0] * (3 - dim)))
            cp2 = (ctypes.c_int32 * 3)(*([int(x) for x in self.cps2] + [1] * (3 - dim)))
            _lib.check(L.sph_nl_fill_max(dim, n, _lib.ptr(r), cs2, cp2, _lib.ptr(counts),
                                         _lib.ptr(info), st), "sph_nl_fill_max")
            max_fill = int(info[2])
            cap = int(max_fill * self.capacity_multiplier) + int(extra_capacity)
            width = min(3 ** dim * max(cap, 1), n)
        else:
            width = n
        if ell_width is not None:
            width = max(1, min(width, int(ell_width)))
        ws = _lib.Workspace(self._config(n, dim, cap, width))
        if ell_width is not None:
            ws.tune("ell_overflow", 1)
        ws.n, ws.dim, ws.cell_capacity = n, dim, cap
        ws.candidate_width = n * (3 ** dim) * cap if self.use_cells else n * n
        bank = 0
        ws.generation[bank] += 1
        _lib.check(L.sph_nl_build(ws.handle, bank, _lib.ptr(r), None, -1, _lib.ptr(info), st),
                   "sph_nl_build")
        max_occ = self._max_occupancy(int(info[1]), ws, extra_capacity)
        _lib.check(L.sph_nl_finalize(ws.handle, max_occ, _lib.ptr(info), st), "sph_nl_finalize")
        self._scan_bits(ws, bank, info)
        self._enable_skin(ws, int(info[3]))
        return NeighborList(self, ws, bank, info, r.clone(), cap, max_occ, self.update)

    def _enable_skin(self, ws, max_pairs):
        """Skin lists for the fused step, sized from the largest exact list."""
        if (self.skin <= 0.0 or not self.use_cells or self.format is not Sparse
                or self.backend != "jaxmd_vmap" or self.dr_threshold != 0.0):
            return
        if self._last_step_info is not None:   # re-allocation: keep what was seen
            self._skin_hint = max(self._skin_hint, int(self._last_step_info[7]))
            self._last_step_info = None
        grow = ((float(self.cutoff) + self.skin) / float(self.cutoff)) ** ws.dim
        width = int(np.ceil(max_pairs * grow * 1.3)) + 8
        width = max(width, int(np.ceil(self._skin_hint * 1.2)) + 8)
        if self._skin_width_override is not None:   # tests: force a small buffer
            width = int(self._skin_width_override)
        ws.enable_skin(self.skin, min(width, ws.n))

    _skin_width_override = None

    def _max_occupancy(self, occ, ws, extra_capacity):
        n = ws.n
        extra = n * int(extra_capacity) if is_sparse(self.format) else int(extra_capacity)
        m = int(_f32(_f32(_f32(occ) * _f32(self.capacity_multiplier)) + _f32(extra)))
        width = ws.candidate_width
        if self.backend == "jaxmd_scan":
            width = self.num_partitions * self._scan_budget(ws)[1]
        m = min(m, width)
        if self.format is OrderedSparse:
            limit = n * (n - 1) // 2
        else:
            limit = n * (n - 1) if self.mask_self else n * n
        return min(m, limit)

    def _scan_budget(self, ws):
        # jax_sph/partition.py:213-235
        n_sub = int(np.ceil(ws.n / self.num_partitions))
        vol = 0.34907 if ws.dim == 2 else 0.15514
        considered = (3 ** ws.dim) * ws.cell_capacity
        return n_sub, int(n_sub * considered * vol * self.capacity_multiplier)

    def _scan_bits(self, ws, bank, info):
        if self.backend != "jaxmd_scan":
            return
        n_sub, budget = self._scan_budget(ws)
        _lib.check(_lib.lib().sph_nl_scan_budget(ws.handle, bank, n_sub, budget,
                                                 _lib.ptr(info), _lib.stream()),
                   "sph_nl_scan_budget")

    def update(self, position, neighbors: NeighborList, **kwargs) -> NeighborList:
        if neighbors._search is not self:
            return neighbors._search.update(position, neighbors, **kwargs)
        ws = neighbors.workspace
        r = _lib.as_f32(position)
        if tuple(r.shape) != (ws.n, ws.dim):
            raise ValueError(f"update got positions of shape {tuple(r.shape)}, the list was "
                             f"allocated for {(ws.n, ws.dim)}")
        L = _lib.lib()
        st = _lib.stream()
        if self.dr_threshold > 0.0:                                  # partition.py:1010-1018
            _lib.check(L.sph_nl_needs_rebuild(ws.handle, _lib.ptr(r),
                                              _lib.ptr(neighbors.reference_position),
                                              float(self.threshold_sq),
                                              _lib.ptr(neighbors.info), st),
                       "sph_nl_needs_rebuild")
            if not int(neighbors.info[4]):
                return neighbors
        info = torch.empty(8, dtype=torch.int64, device=r.device)
        bank = self.next_exact_bank(neighbors)
        ws.generation[bank] += 1
        _lib.check(L.sph_nl_build(ws.handle, bank, _lib.ptr(r), _lib.ptr(neighbors.info),
                                  neighbors.max_occupancy, _lib.ptr(info), st), "sph_nl_build")
        self._scan_bits(ws, bank, info)
        return NeighborList(self, ws, bank, info, r.clone() if self.dr_threshold > 0 else r,
                            neighbors.cell_list_capacity, neighbors.max_occupancy,
                            neighbors.update_fn)

    @staticmethod
    def next_exact_bank(neighbors: NeighborList) -> int:
        """Bank for the next exact list: the one ``neighbors`` does not use."""
        return 1 - neighbors.bank if neighbors.bank < 2 else 0

    def step_bank(self, neighbors: NeighborList) -> int:
        """Bank the fused step writes: the skin lists when enabled."""
        ws = neighbors.workspace
        return _lib.SKIN_BANK if ws.skin_on else self.next_exact_bank(neighbors)

    def adopt_step(self, neighbors: NeighborList, bank: int, info: torch.Tensor,
                   r: torch.Tensor) -> NeighborList:
        """NeighborList for a bank just written by the fused step."""
        if bank == _lib.SKIN_BANK:
            self._last_step_info = info
        else:
            self._scan_bits(neighbors.workspace, bank, info)
        return NeighborList(self, neighbors.workspace, bank, info, r,
                            neighbors.cell_list_capacity, neighbors.max_occupancy,
                            neighbors.update_fn)

    def write_idx(self, nbrs: NeighborList) -> torch.Tensor:
        m = nbrs.max_occupancy
        idx = torch.empty((2, m), dtype=torch.int32, device=nbrs.info.device)
        if m and nbrs.bank == _lib.SKIN_BANK:
            r = _lib.as_f32(nbrs.reference_position)
            _lib.check(_lib.lib().sph_nl_write_idx_skin(nbrs.workspace.handle, _lib.ptr(r),
                                                        _lib.ptr(idx), m, _lib.stream()),
                       "sph_nl_write_idx_skin")
        elif m:
            _lib.check(_lib.lib().sph_nl_write_idx(nbrs.workspace.handle, nbrs.bank, _lib.ptr(idx),
                                                   m, _lib.ptr(nbrs.info), _lib.stream()),
                       "sph_nl_write_idx")
        return idx

    def exact_at_reference(self, nbrs: NeighborList) -> NeighborList:
        """An exact list (bank 0/1) of the pairs of a skin list, built at its
        reference position; used when a skin list meets other positions."""
        ws = nbrs.workspace
        info = torch.empty(8, dtype=torch.int64, device=nbrs.info.device)
        bank = self.next_exact_bank(nbrs)
        ws.generation[bank] += 1
        r = _lib.as_f32(nbrs.reference_position)
        _lib.check(_lib.lib().sph_nl_build(ws.handle, bank, _lib.ptr(r), _lib.ptr(nbrs.info),
                                           nbrs.max_occupancy, _lib.ptr(info), _lib.stream()),
                   "sph_nl_build")
        return NeighborList(self, ws, bank, info, r, nbrs.cell_list_capacity,
                            nbrs.max_occupancy, nbrs.update_fn)


def neighbor_list(displacement_or_metric, box, r_cutoff: float, dr_threshold: float = 0.0,
                  capacity_multiplier: float = 1.25, disable_cell_list: bool = False,
                  mask_self: bool = True, custom_mask_function=None,
                  fractional_coordinates: bool = False, format: NeighborListFormat = Dense,
                  **static_kwargs) -> NeighborListFns:
    """partition.py:723-1026 (vmap backend semantics)."""
    s = _NeighborSearch(displacement_or_metric, box, r_cutoff, dr_threshold,
                        capacity_multiplier, disable_cell_list, mask_self,
                        custom_mask_function, fractional_coordinates, format,
                        backend=static_kwargs.get("backend", "jaxmd_vmap"),
                        num_partitions=static_kwargs.get("num_partitions", 1),
                        skin=static_kwargs.get("skin", "auto"))

    def allocate_fn(position, extra_capacity: int = 0, **kwargs):
        return s.allocate(position, extra_capacity, **kwargs)

    def update_fn(position, neighbors, **kwargs):
        return s.update(position, neighbors, **kwargs)

    fns = NeighborListFns(allocate_fn, update_fn)
    fns.search = s
    return fns


# ------------------------------------------------------------------ cell lists
class CellList:
    """partition.py:53-101.  Buffers have the common leading shape
    ``S = (nz, ny, nx, cell_capacity)`` (``(ny, nx, cell_capacity)`` in 2D):
    ``position_buffer`` S + (dim,) float32, ``id_buffer`` S + (1,) int32 with
    empty slots = N, ``named_buffer[k]`` S + the side array's trailing shape
    (``(1,)`` for a 1-d array) with empty slots = 10**5.
    ``did_buffer_overflow`` is a 0-d bool device tensor."""

    def __init__(self, position_buffer, id_buffer, named_buffer, did_buffer_overflow,
                 cell_capacity: int, cell_size, update_fn: Callable):
        self.position_buffer = position_buffer
        self.id_buffer = id_buffer
        self.named_buffer = named_buffer
        self.did_buffer_overflow = did_buffer_overflow
        self.cell_capacity = cell_capacity
        self.cell_size = cell_size
        self.update_fn = update_fn

    def update(self, position, **kwargs) -> "CellList":
        cl_data = (self.cell_capacity, self.did_buffer_overflow, self.update_fn)
        return self.update_fn(position, cl_data, **kwargs)

    @property
    def kwarg_buffers(self):
        return self.named_buffer


class CellListFns:
    """partition.py:104-111."""

    def __init__(self, allocate: Callable, update: Callable):
        self.allocate = allocate
        self.update = update

    def __iter__(self):
        return iter((self.allocate, self.update))


class _CellGrid:
    """_cell_dimensions (partition.py:112-156).  A Python-scalar box is
    divided in float64 and gives ``(cps,) * dim`` cells with no size check, as
    in the reference; an array box is divided in float32 (this package's
    precision) and must hold at least 3 cells per side."""

    def __init__(self, dim: int, box_size, minimum_cell_size):
        if isinstance(box_size, torch.Tensor):
            box_size = box_size.detach().cpu().numpy()
        self.scalar = isinstance(box_size, (int, float, np.integer, np.floating))
        if self.scalar:
            box = float(box_size)
            cps = np.floor(box / float(minimum_cell_size))
            cell_size = box / cps
            self.cps = np.full(dim, int(cps), np.int64)
            self.cell_size = cell_size                     # float, as the reference stores it
            self.cell_size32 = np.full(dim, np.float32(cell_size), np.float32)
            self.shape = (int(cps),) * dim
        else:
            b = np.asarray(box_size)
            box = b.astype(np.float32)
            if box.ndim == 0:
                cps = np.floor(box / np.float32(minimum_cell_size))
                self.cell_size = (box / cps).astype(np.float32)
                self.cps = np.full(dim, int(cps), np.int64)
                self.cell_size32 = np.full(dim, self.cell_size, np.float32)
                self.shape = (int(cps),) * dim
            else:
                box = box.reshape(-1)
                assert box.size == dim
                mcs = np.asarray(minimum_cell_size, np.float32)
                cps = np.floor(box / mcs)
                self.cell_size = (box / cps).astype(np.float32)
                self.cps = cps.astype(np.int64)
                if np.any(self.cps < 3):
                    raise ValueError("Box must be at least 3x the size of the grid spacing in "
                                     "each dimension.")
                self.cell_size32 = self.cell_size.reshape(-1)
                self.shape = tuple(int(x) for x in self.cps[::-1])
        if np.any(self.cps < 1):
            raise ValueError(f"Box {box_size} holds no cell of size {minimum_cell_size}.")
        self.box_size = box_size
        self.count = int(np.prod(self.cps))

    def c_args(self, dim):
        cs = (ctypes.c_float * 3)(*([float(x) for x in self.cell_size32] + [1.0] * (3 - dim)))
        cp = (ctypes.c_int32 * 3)(*([int(x) for x in self.cps] + [1] * (3 - dim)))
        return cs, cp


def _cell_counts(position: torch.Tensor, grid: _CellGrid, info: torch.Tensor) -> torch.Tensor:
    n, dim = position.shape
    counts = torch.empty(grid.count, dtype=torch.int32, device=position.device)
    cs, cp = grid.c_args(dim)
    r = position if n else torch.zeros(1, dtype=torch.float32, device=position.device)
    _lib.check(_lib.lib().sph_nl_fill_max(dim, n, _lib.ptr(r), cs, cp, _lib.ptr(counts),
                                          _lib.ptr(info), _lib.stream()), "sph_nl_fill_max")
    return counts


def count_cell_filling(position, box_size, minimum_cell_size) -> torch.Tensor:
    """partition.py:159-174: particles per cell, flat over the cell hash
    (int32 on the device).  A particle on the upper box face counts in the
    last cell (the reference hashes it into a wrong one)."""
    r = _lib.as_f32(position)
    grid = _CellGrid(int(r.shape[1]), box_size, minimum_cell_size)
    info = torch.zeros(8, dtype=torch.int64, device=r.device)
    return _cell_counts(r, grid, info)


_EMPTY_KWARG_VALUE = 10 ** 5  # partition.py:334


def _side_array(k, v, n, dev):
    """A per-particle side array as 32-bit words (partition.py:336-357):
    floating types in float32, integer types in int32 (this package's
    precision), bool through int32."""
    if not isinstance(v, (torch.Tensor, np.ndarray)):
        raise ValueError(f'Data must be specified as an ndarray. Found "{k}" with type {type(v)}.')
    if v.shape[0] != n:
        raise ValueError("Data must be specified per-particle (an ndarray with shape "
                         f'({n}, ...)). Found "{k}" with shape {tuple(v.shape)}.')
    t = torch.as_tensor(v)
    if t.dtype == torch.bool:
        out_dtype, t = torch.bool, t.to(torch.int32)
    elif t.is_floating_point():
        t = t.to(torch.float32)
        out_dtype = torch.float32
    else:
        t = t.to(torch.int32)
        out_dtype = torch.int32
    tail = tuple(v.shape[1:]) if len(v.shape) > 1 else (1,)
    t = t.to(dev).contiguous()
    fill = (np.array([_EMPTY_KWARG_VALUE], np.float32).view(np.uint32)[0]
            if t.dtype == torch.float32 else _EMPTY_KWARG_VALUE)
    return t, tail, out_dtype, int(fill)


def cell_list(box_size, minimum_cell_size: float,
              buffer_size_multiplier: float = 1.25) -> CellListFns:
    """partition.py:243-428 on the HIP library (``sph_cell_list``): hashing,
    the stable cell sort, the slot scatter and the occupancy all run on the
    device; ``allocate`` reads the largest cell occupancy back to size the
    capacity, ``update`` never synchronises."""

    def cell_list_fn(position, capacity_overflow_update=None, extra_capacity: int = 0,
                     **kwargs) -> CellList:
        dev = _lib.device()
        r = _lib.as_f32(position, dev)
        if r.dim() != 2 or r.shape[1] not in (2, 3):
            raise ValueError(f"Cell list spatial dimension must be 2 or 3. Found "
                             f"{r.shape[-1] if r.dim() else 0}.")
        n, dim = r.shape
        grid = _CellGrid(dim, box_size, minimum_cell_size)
        info = torch.zeros(8, dtype=torch.int64, device=dev)
        if capacity_overflow_update is None:
            # _estimate_cell_capacity (partition.py:195-199) on the grid
            # re-derived from the rounded cell size
            grid2 = _CellGrid(dim, grid.box_size, grid.cell_size)
            _cell_counts(r, grid2, info)
            cell_capacity = int(int(info[2]) * buffer_size_multiplier) + int(extra_capacity)
            overflow = torch.zeros((), dtype=torch.bool, device=dev)
            update_fn = cell_list_fn
        else:
            cell_capacity, overflow, update_fn = capacity_overflow_update
            overflow = torch.as_tensor(overflow, device=dev).to(torch.bool)
        cell_capacity = int(cell_capacity)
        side = [(k,) + _side_array(k, v, n, dev) for k, v in kwargs.items()]
        slots = grid.count * cell_capacity
        pos_buf = torch.empty(slots * dim, dtype=torch.float32, device=dev)
        id_buf = torch.empty(slots, dtype=torch.int32, device=dev)
        named = {}
        for k, t, tail, _, _ in side:
            named[k] = torch.empty(slots * int(np.prod(tail)), dtype=t.dtype, device=dev)
        if slots == 0:
            pass  # no slot to write: only an empty particle set sizes a capacity of 0
        else:
            L = _lib.lib()
            work = torch.empty(max(1, L.sph_cell_list_work_bytes(n, grid.count)),
                               dtype=torch.uint8, device=dev)
            cs, cp = grid.c_args(dim)
            k = len(side)
            src = (ctypes.c_void_p * max(k, 1))(*[t.data_ptr() for _, t, *_ in side])
            dst = (ctypes.c_void_p * max(k, 1))(*[named[s[0]].data_ptr() for s in side])
            width = (ctypes.c_int32 * max(k, 1))(*[int(np.prod(s[2])) for s in side])
            fill = (ctypes.c_uint32 * max(k, 1))(*[s[4] for s in side])
            rr = r if n else torch.zeros(1, dtype=torch.float32, device=dev)
            _lib.check(L.sph_cell_list(dim, n, _lib.ptr(rr), cs, cp, cell_capacity,
                                       _lib.ptr(pos_buf), _lib.ptr(id_buf), k, src, width, fill,
                                       dst, _lib.ptr(work), work.numel(), _lib.ptr(info),
                                       _lib.stream()), "sph_cell_list")
            overflow = overflow | ((info[0] & int(PEC.CELL_LIST_OVERFLOW)) != 0)
        shape = grid.shape + (cell_capacity,)
        named_out = {}
        for k, t, tail, out_dtype, _ in side:
            b = named[k].view(shape + tail)
            named_out[k] = b != 0 if out_dtype == torch.bool else b
        return CellList(pos_buf.view(shape + (dim,)), id_buf.view(shape + (1,)), named_out,
                        overflow, cell_capacity, grid.cell_size, update_fn)

    def allocate_fn(position, extra_capacity: int = 0, **kwargs) -> CellList:
        return cell_list_fn(position, extra_capacity=extra_capacity, **kwargs)

    def update_fn(position, cl_or_capacity, **kwargs) -> CellList:
        if isinstance(cl_or_capacity, int):
            return cell_list_fn(position, (int(cl_or_capacity), False, cell_list_fn), **kwargs)
        cl = cl_or_capacity
        return cell_list_fn(position, (cl.cell_capacity, cl.did_buffer_overflow, cl.update_fn),
                            **kwargs)

    return CellListFns(allocate_fn, update_fn)

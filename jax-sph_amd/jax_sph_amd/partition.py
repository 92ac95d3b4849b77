"""Neighbour-search backends (drop-in for ``jax_sph/partition.py:485-571``).

``jaxmd_vmap`` and ``jaxmd_scan`` both run on the HIP cell-list search and give
the same pair set; ``jaxmd_scan`` additionally reproduces the chunked edge
budget of ``_scan_neighbor_list`` (``partition.py:213-291``): a chunk of
``ceil(N / num_partitions)`` senders holding more pairs than its budget sets
``CELL_LIST_OVERFLOW``, and ``max_occupancy`` is capped by the chunk buffers.
The ``matscipy`` backend (a CPU callback) is not provided.
"""
from __future__ import annotations

from .jax_md import space  # noqa: F401
from .jax_md.partition import (  # noqa: F401
    CellList,
    CellListFns,
    Dense,
    NeighborList,
    NeighborListFns,
    NeighborListFormat,
    OrderedSparse,
    PartitionError,
    PartitionErrorCode,
    Sparse,
    StaleNeighborListError,
    cell_list,
    count_cell_filling,
    is_sparse,
)
from .jax_md.partition import neighbor_list as vmap_neighbor_list

PEC = PartitionErrorCode


def _scan_neighbor_list(displacement_or_metric, box, r_cutoff, dr_threshold=0.0,
                        capacity_multiplier=1.25, disable_cell_list=False, mask_self=True,
                        custom_mask_function=None, fractional_coordinates=False,
                        format=NeighborListFormat.Sparse, num_partitions=8, **static_kwargs):
    """partition.py:55-353 semantics on the HIP search."""
    assert disable_cell_list is False, "Works only with a cell list"
    assert not fractional_coordinates, "Works only with real coordinates"
    assert format == NeighborListFormat.Sparse, "Works only with sparse neighbor list"
    assert custom_mask_function is None, "Custom masking not implemented"
    return vmap_neighbor_list(displacement_or_metric, box, r_cutoff, dr_threshold,
                              capacity_multiplier, disable_cell_list, mask_self, None, False,
                              format, backend="jaxmd_scan", num_partitions=num_partitions)


def _vmap_neighbor_list(displacement_or_metric, box, r_cutoff, dr_threshold=0.0,
                        capacity_multiplier=1.25, disable_cell_list=False, mask_self=True,
                        custom_mask_function=None, fractional_coordinates=False,
                        format=NeighborListFormat.Sparse, skin="auto", **static_kwargs):
    return vmap_neighbor_list(displacement_or_metric, box, r_cutoff, dr_threshold,
                              capacity_multiplier, disable_cell_list, mask_self,
                              custom_mask_function, fractional_coordinates, format, skin=skin)


def _matscipy_neighbor_list(*args, **kwargs):
    raise NotImplementedError(
        "the matscipy backend is a CPU callback (partition.py:356-482) and is not part of "
        "this HIP implementation; use backend='jaxmd_vmap'")


BACKENDS = {
    "jaxmd_vmap": _vmap_neighbor_list,
    "jaxmd_scan": _scan_neighbor_list,
    "matscipy": _matscipy_neighbor_list,
}


def neighbor_list(displacement_or_metric, box_size, r_cutoff: float,
                  backend: str = "jaxmd_vmap", dr_threshold: float = 0.0,
                  capacity_multiplier: float = 1.25, disable_cell_list: bool = False,
                  mask_self: bool = True, custom_mask_function=None,
                  fractional_coordinates: bool = False,
                  format: NeighborListFormat = NeighborListFormat.Sparse,
                  num_particles_max: int = None, num_partitions: int = 1, pbc=None,
                  skin="auto"):
    """partition.py:492-571: same arguments; ``num_particles_max``/``pbc`` only
    matter to the matscipy backend.  ``skin`` (not in the reference) sets the
    Verlet skin of the fused step's lists for ``jaxmd_vmap`` ("auto" = 0.08 *
    cutoff, 0 = rebuild exact lists every step); results are the same."""
    assert backend in BACKENDS, f"Unknown backend {backend}"
    kw = dict(num_particles_max=num_particles_max, num_partitions=num_partitions, pbc=pbc)
    if backend == "jaxmd_vmap":
        kw["skin"] = skin
    return BACKENDS[backend](
        displacement_or_metric, box_size, r_cutoff, dr_threshold, capacity_multiplier,
        disable_cell_list, mask_self, custom_mask_function, fractional_coordinates, format, **kw)

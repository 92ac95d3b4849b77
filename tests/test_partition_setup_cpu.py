"""Set-up errors of ``partition.neighbor_list``, raised on the host before any
device work, with the exception types of the reference:
* an unknown backend (partition.py:555) and the scan backend's restrictions
  (partition.py:149-152) are AssertionErrors;
* a format that is not a ``NeighborListFormat`` is a ValueError
  (jax_md/partition.py:603-609);
* the matscipy backend, Dense lists, custom masks and fractional coordinates
  of the vmap backend are not implemented here and say so.
The few-cells assertion of the scan backend (partition.py:167) needs the
positions, so it is checked with ``allocate`` in the GPU tests."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "jax-sph_amd"))

from jax_sph_amd import partition  # noqa: E402
from jax_sph_amd.jax_md import space  # noqa: E402

BOX = np.array([1.0, 1.0])
DISP, _ = space.periodic(side=BOX)
SPARSE = partition.NeighborListFormat.Sparse


def _make(**kw):
    args = dict(r_cutoff=0.1, format=SPARSE)
    args.update(kw)
    return partition.neighbor_list(DISP, BOX, **args)


def test_valid_configurations_construct_without_a_device():
    assert _make(backend="jaxmd_vmap").search.backend == "jaxmd_vmap"
    fns = _make(backend="jaxmd_scan", num_partitions=4)
    assert fns.search.backend == "jaxmd_scan" and fns.search.num_partitions == 4
    assert _make(skin=0.0).search.skin == 0.0
    assert _make(skin="auto").search.skin == pytest.approx(0.008, rel=1e-6)


def test_unknown_backend():
    with pytest.raises(AssertionError, match="Unknown backend"):
        _make(backend="cell_magic")


@pytest.mark.parametrize("kw,msg", [
    ({"disable_cell_list": True}, "Works only with a cell list"),
    ({"fractional_coordinates": True}, "Works only with real coordinates"),
    ({"format": partition.NeighborListFormat.Dense}, "Works only with sparse neighbor list"),
    ({"custom_mask_function": lambda idx: idx}, "Custom masking not implemented"),
])
def test_scan_backend_restrictions(kw, msg):
    with pytest.raises(AssertionError, match=msg):
        _make(backend="jaxmd_scan", **kw)


def test_format_must_be_a_neighbor_list_format():
    with pytest.raises(ValueError, match="must be a member of NeighborListFormat"):
        _make(format="sparse")


@pytest.mark.parametrize("kw", [
    {"backend": "matscipy"},
    {"format": partition.NeighborListFormat.Dense},
    {"custom_mask_function": lambda idx: idx},
    {"fractional_coordinates": True},
])
def test_unsupported_options_say_so(kw):
    with pytest.raises(NotImplementedError):
        _make(**kw)


def test_negative_skin():
    with pytest.raises(ValueError, match="skin"):
        _make(skin=-0.01)

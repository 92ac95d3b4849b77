"""Pins the oracle's dense cell list (oracle/nl_oracle.py ``cell_list_build``,
restating jax_md/partition.py:294-413) through a second formulation of its
slot layout, brute-force cell membership and the capacity rule; the GPU tests
of ``cell_list`` compare the library with it.

The reference scatters the particle at stable-argsort position s to slot
``hash * capacity + s mod capacity``.  Written per cell, that is: the cell's
particles, in ascending id, fill the slots cyclically from ``(number of
particles in lower cells) mod capacity``; a cell with more particles than
slots wraps around and the later particle keeps a shared slot."""
import numpy as np
import pytest

from oracle import nl_oracle as nlo


def _cells_by_brute_force(r, cell_size, cps):
    """Flat cell of every particle: floor of the float32 quotient per axis,
    clamped into the grid, hashed x + nx*(y + ny*z)."""
    n, dim = r.shape
    out = np.zeros(n, np.int64)
    for i in range(n):
        h, m = 0, 1
        for d in range(dim):
            q = int(np.floor(np.float32(r[i, d]) / np.float32(cell_size[d])))
            q = min(max(q, 0), int(cps[d]) - 1)
            h += q * m
            m *= int(cps[d])
        out[i] = h
    return out


def _expected_slots(cell, cell_count, cap, n):
    """The per-cell cyclic fill, built cell by cell with plain loops."""
    want = np.full((cell_count, cap), n, np.int64)
    start = 0
    for c in range(cell_count):
        members = [i for i in range(n) if cell[i] == c]   # ascending id
        for j, i in enumerate(members):
            want[c, (start + j) % cap] = i                # a later member overwrites
        start += len(members)
    return want


@pytest.mark.parametrize("dim,n,box,mcs,seed", [
    (2, 300, [1.0, 1.0], 0.1, 0), (2, 257, [1.3, 0.7], 0.12, 1),
    (3, 400, [1.0, 1.0, 1.0], 0.2, 2), (3, 333, [0.9, 1.2, 1.0], 0.25, 3)])
def test_slots_fill_each_cell_cyclically(dim, n, box, mcs, seed):
    rng = np.random.default_rng(seed)
    box = np.array(box, np.float32)
    r = (rng.random((n, dim)) * box).astype(np.float32)
    cl = nlo.cell_list_build(r, box, np.float32(mcs), clamp=True)
    cell = _cells_by_brute_force(r, cl.cell_size, cl.cps)
    np.testing.assert_array_equal(cl.particle_cell, cell)
    assert not cl.overflow
    want = _expected_slots(cell, cl.id_buffer.shape[0], cl.capacity, n)
    np.testing.assert_array_equal(cl.id_buffer, want)
    # every particle once, in its own cell
    ids = cl.id_buffer[cl.id_buffer < n]
    np.testing.assert_array_equal(np.sort(ids), np.arange(n))


def test_capacity_is_the_largest_filling_times_the_multiplier():
    """_estimate_cell_capacity (partition.py:195-199) plus extra capacity, the
    filling counted on the grid re-derived from the rounded cell size."""
    rng = np.random.default_rng(7)
    box = np.array([1.0, 1.0], np.float32)
    r = (rng.random((500, 2)) * box).astype(np.float32)
    for mult, extra in [(1.25, 0), (1.0, 0), (2.0, 3)]:
        cl = nlo.cell_list_build(r, box, np.float32(0.1), buffer_size_multiplier=mult,
                                 extra_capacity=extra, clamp=True)
        cs, cps, _ = nlo.cell_dimensions(box, cl.cell_size)
        fill = np.bincount(_cells_by_brute_force(r, cs, cps), minlength=int(np.prod(cps)))
        assert cl.capacity == int(int(fill.max()) * mult) + extra


def test_overflowing_cell_wraps_and_the_later_particle_keeps_the_slot():
    """Five particles in one cell of capacity 3 (one earlier particle shifts
    the start): slots hold the last writer; overflow is flagged."""
    box = np.array([1.0, 1.0], np.float32)
    r = np.array([[0.05, 0.05]] + [[0.45 + 0.01 * k, 0.45] for k in range(5)], np.float32)
    cl = nlo.cell_list_build(r, box, np.float32(0.1), capacity=3, clamp=True)
    assert cl.overflow
    c = int(cl.particle_cell[1])
    # cell 0 holds particle 0 at slot 0; the big cell starts at sorted position 1
    assert cl.id_buffer[0, 0] == 0
    # positions 1..5 -> slots 1,2,0,1,2: slot 0 <- particle 3, 1 <- 4, 2 <- 5
    np.testing.assert_array_equal(cl.id_buffer[c], [3, 4, 5])
    np.testing.assert_array_equal(cl.id_buffer, _expected_slots(cl.particle_cell,
                                                                cl.id_buffer.shape[0], 3, 6))


def test_upper_face_is_clamped_into_the_last_cell():
    """The library bins r == side into the last cell; the reference's int32
    hash would put it one past the grid (the oracle raises unless clamped)."""
    box = np.array([1.0, 1.0], np.float32)
    r = np.array([[1.0, 0.5], [0.5, 0.5], [0.95, 0.5]], np.float32)
    with pytest.raises(nlo.EdgeOfBoxError):
        nlo.cell_list_build(r, box, np.float32(0.1))
    cl = nlo.cell_list_build(r, box, np.float32(0.1), clamp=True)
    assert cl.particle_cell[0] == cl.particle_cell[2]

"""sph_nl_fill_max_static, the cell-occupancy count that allocates nothing
(its cell counts live in a static device array), against sph_nl_fill_max with
caller scratch: two calls in a row on one GPU, the second on a smaller grid
and fewer particles, so the second call also shows that the shared array is
cleared between calls."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

js = pytest.importorskip("jax_sph_amd")
from jax_sph_amd import _lib  # noqa: E402

SPH_ERR_SIZE = -3


def _case(dim, n, cps, seed):
    g = np.random.default_rng(seed)
    cps = np.asarray(cps, np.int32)
    box = np.float32(1.0)
    cell = (box / cps).astype(np.float32)
    # clustered positions, so one cell holds clearly more particles than the rest
    r = g.random((n, dim), dtype=np.float32) * box
    r[: n // 10] = (g.random((n // 10, dim), dtype=np.float32) * 0.02 + 0.5).astype(np.float32)
    return r.astype(np.float32), cell, cps


def _host_max(r, cell, cps):
    q = np.floor((r.astype(np.float64) / cell.astype(np.float64)).astype(np.float32)).astype(
        np.int64)
    q = np.clip(q, 0, cps.astype(np.int64) - 1)
    h = q[:, 0] + q[:, 1] * cps[0]
    if r.shape[1] == 3:
        h = h + q[:, 2] * cps[0] * cps[1]
    return int(np.bincount(h, minlength=int(np.prod(cps.astype(np.int64)))).max())


def _call(static, r_np, cell, cps):
    L = _lib.lib()
    st = _lib.stream()
    n, dim = r_np.shape
    r = torch.from_numpy(r_np).cuda()
    info = torch.full((8,), -7, dtype=torch.int64, device="cuda")
    cs3 = (ctypes.c_float * 3)(*([float(x) for x in cell] + [1.0] * (3 - dim)))
    cp3 = (ctypes.c_int32 * 3)(*([int(x) for x in cps] + [1] * (3 - dim)))
    if static:
        rc = L.sph_nl_fill_max_static(dim, n, _lib.ptr(r), cs3, cp3, _lib.ptr(info), st)
    else:
        counts = torch.empty(int(np.prod(cps.astype(np.int64))), dtype=torch.int32,
                             device="cuda")
        rc = L.sph_nl_fill_max(dim, n, _lib.ptr(r), cs3, cp3, _lib.ptr(counts), _lib.ptr(info),
                               st)
    _lib.check(rc, "fill_max")
    torch.cuda.synchronize()
    return info.cpu().numpy()


def test_static_scratch_fill_max_matches_caller_scratch_twice():
    cases = [(3, 200_000, (64, 64, 64), 0), (2, 3_000, (10, 10), 1)]  # second one smaller
    for dim, n, cps, seed in cases:
        r, cell, cps = _case(dim, n, cps, seed)
        print(f"fill_max dim={dim} n={n} cps={tuple(cps)}", flush=True)
        got = _call(True, r, cell, cps)
        ref = _call(False, r, cell, cps)
        assert got[2] == ref[2] == _host_max(r, cell, cps), (got[2], ref[2])
        # only info[2] is written
        assert (np.delete(got, 2) == -7).all() and (np.delete(ref, 2) == -7).all()


def test_static_scratch_rejects_grids_beyond_its_size():
    L = _lib.lib()
    cps = (256, 256, 256)  # 2^24 cells > SPH_FILL_MAX_STATIC_CELLS = 2^21
    r = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    info = torch.zeros(8, dtype=torch.int64, device="cuda")
    cs3 = (ctypes.c_float * 3)(*([1.0 / 256] * 3))
    cp3 = (ctypes.c_int32 * 3)(*cps)
    rc = L.sph_nl_fill_max_static(3, 4, _lib.ptr(r), cs3, cp3, _lib.ptr(info), _lib.stream())
    assert rc == SPH_ERR_SIZE
    assert b"static scratch" in L.sph_last_error()

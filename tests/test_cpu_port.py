"""The OpenMP C restatement of the step (oracle/sph_cpu.c, the benchmark's
multi-core CPU baseline) against the numpy oracle: the same rollouts at the
float32 noise floor of the float64 oracle, in 2D and 3D, with and without the
transport-velocity formulation, on 1 and several threads (bitwise equal: each
particle gathers its own sums)."""
import numpy as np
import pytest

from oracle import cpu_port
from oracle import nl_oracle as nlo
from oracle import sph_oracle as so
from tests import helpers as H
from tests import gpu_helpers as G

pytestmark = pytest.mark.skipif(
    not __import__("os").path.exists(cpu_port.LIB), reason="oracle/libsph_cpu.so not built")


def _tgv(dim, n, dtype=np.float32):
    side = 2 * np.pi if dim == 3 else 1.0
    dx = side / n
    r = H.lattice(n, dx, dim, dtype)
    st = H.base_state(r, dx=dx, eta=0.01, dtype=dtype)
    x, y = r[:, 0] * (2 * np.pi / side), r[:, 1] * (2 * np.pi / side)
    u = np.zeros_like(r)
    if dim == 2:
        u[:, 0] = -np.cos(x) * np.sin(y)
        u[:, 1] = np.sin(x) * np.cos(y)
    else:
        z = r[:, 2]
        u[:, 0] = np.sin(x) * np.cos(y) * np.cos(z)
        u[:, 1] = -np.cos(x) * np.sin(y) * np.cos(z)
    st["u"] = u.astype(dtype)
    st["v"] = st["u"].copy()
    return st, np.full(dim, side), dx


@pytest.mark.parametrize("dim,n,kernel,tvf,steps", [
    (2, 20, "QSK", 0.0, 6), (2, 20, "WC2K", 1.0, 6), (3, 10, "QSK", 1.0, 4),
])
def test_cpu_port_matches_oracle(dim, n, kernel, tvf, steps):
    st32, box, dx = _tgv(dim, n)
    h = dx * (1.3 if kernel == "WC2K" else 1.0)
    dt = 0.25 * h / 11.0
    solvers = {}
    for t in (np.float32, np.float64):
        eos = so.TaitEoS(100.0, 1.0, 0.0, 1.0, t)
        solvers[t] = so.WCSPH(box, eos, lambda q: np.zeros_like(q), dx, dim, dt, 10.0,
                              kernel=kernel, h_fac=h / dx, dtype=t)
    runs = {}
    for t in (np.float32, np.float64):
        s = solvers[t]
        ns = nlo.NeighborSearch(box, s.kernel.cutoff, mask_self=False, dtype=t)
        adv = so.si_euler(tvf, s.forward, box, None, ns.update, t)
        st = G.to_f64(dict(st32)) if t is np.float64 else dict(st32)
        nb = ns.allocate(st["r"])
        for _ in range(steps):
            st, nb = adv(dt, st, nb)
        runs[t] = st
    s32 = solvers[np.float32]
    out = {}
    for threads in (1, 4):
        st = {k: np.array(v, copy=True) for k, v in st32.items()}
        step = cpu_port.CPUStep(box, s32.kernel, s32.eos, dt, tvf, threads=threads)
        for _ in range(steps):
            step(st)
        out[threads] = st
    for k in ("r", "u", "v", "rho", "p", "dudt", "dvdt"):
        assert np.array_equal(out[1][k], out[4][k]), k
    G.assert_within_noise_floor(out[1], runs[np.float32], runs[np.float64], 1e-5,
                                fields=("r", "u", "v", "rho"), label="cpu port")
    G.assert_within_noise_floor(out[1], runs[np.float32], runs[np.float64], 1e-5,
                                fields=("dudt", "dvdt", "p"), factor=8.0, label="cpu port")

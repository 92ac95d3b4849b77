"""Input builders shared by the tests (lattices, states, physics settings)."""
import numpy as np

from oracle import nl_oracle as nlo
from oracle import sph_oracle as so


def lattice(n, dx, dim, dtype=np.float32):
    """Cell-centred lattice fl(fl(i+0.5)*fl(dx)) with the reference's meshgrid order
    (``utils.py:35-53``): 2D k = iy*nx+ix, 3D k = (iy*nx+ix)*nz+iz."""
    ns = [n] * dim if np.isscalar(n) else list(n)
    t = np.dtype(dtype).type
    if dim == 2:
        g = np.meshgrid(range(ns[0]), range(ns[1]), indexing="xy")
    else:
        g = np.meshgrid(range(ns[0]), range(ns[1]), range(ns[2]), indexing="xy")
    idx = np.vstack([x.ravel() for x in g]).T
    return ((idx + t(0.5)).astype(dtype) * t(dx)).astype(dtype)


def base_state(r, rho=1.0, mass=None, eta=0.01, dx=0.02, dtype=np.float32):
    N, dim = r.shape
    t = np.dtype(dtype).type
    m = t(dx ** dim) if mass is None else t(mass)
    z = np.zeros((N, dim), dtype)
    return {
        "r": r.astype(dtype), "tag": np.zeros(N, np.int64), "u": z.copy(), "v": z.copy(),
        "dudt": z.copy(), "dvdt": z.copy(), "drhodt": np.zeros(N, dtype),
        "rho": np.full(N, rho, dtype), "p": np.zeros(N, dtype),
        "mass": np.full(N, m, dtype), "eta": np.full(N, eta, dtype),
        "dTdt": np.zeros(N, dtype), "T": np.ones(N, dtype), "kappa": np.zeros(N, dtype),
        "Cp": np.zeros(N, dtype), "nw": z.copy(),
    }


def tait(c_ref=10.0, rho_ref=1.0, gamma=1.0, p_bg_factor=0.0, dtype=np.float32):
    p_ref = rho_ref * c_ref ** 2 / gamma
    return so.TaitEoS(p_ref, rho_ref, p_bg_factor * p_ref, gamma, dtype)


def zero_g(r):
    return np.zeros_like(r)


def pairs(r, box, cutoff, dtype=np.float32):
    ns = nlo.NeighborSearch(np.asarray(box), cutoff, mask_self=False, dtype=dtype)
    return ns.allocate(r).idx

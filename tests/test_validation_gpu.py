"""Validation of the fused step against published solutions, as the
reference's validation scripts do (validation/tgv2d.sh, tgv3d.sh, ldc2d.sh,
validate.py), from relaxed initial positions where the reference's scripts
relax first:

* 2D Taylor-Green vortex, Re = 100, transport velocity: kinetic energy
  against the analytic decay 0.25 exp(-16 pi^2 t / Re);
* 3D Taylor-Green vortex, Re = 100: kinetic energy per volume against the
  JAX-Fluids solution (Nx = 64) shipped with the reference
  (validation/ref/tgv3d_ref_100.txt, tests/golden/validation_refs.json);
* lid-driven cavity, Re = 100: centreline velocities through
  sph_interpolator against Ghia et al. (1982)
  (validation/ref/ldc_data_{u,v}_vel.csv);
* reverse Poiseuille flow (from the lattice, as its case file starts): the
  velocity profile at t = 0.2, 1, 2 and 10 against the closed-form start-up
  series and the reference's profiles (validation/ref/rpf_t_*.csv), within
  0.12-0.56% and 0.45-1.4% of the peak speed (MI355X);
* dam break of a width-a, height-2a column (validation/db2d.sh's Riemann
  settings): the surge front against Martin & Moyce's experiments
  (validation/ref/dambreak_data.py).

The reference checks these by eye on plots; here each comparison has a bound
set above the error measured at these resolutions, and the errors must fall
as the resolution rises.  Measured (MI355X): 2D TGV -4.2% / -2.3% at t = 0.5
and -8.3% / -4.4% at t = 1 (nx = 50 / 100); 3D TGV largest error over
t <= 10 of 19.7% / 15.5% (nx = 24 / 32), under 3.4% for t <= 2.

The cavity's centrelines at dx = 0.02 carry particle noise: single snapshots
between t = 15 and 20 differ by 0.02-0.05, and the last one moves with the
relaxed arrangement and with summation order.  The test therefore averages
the snapshots at t = 15, 16, ..., 20.  Over relaxation seeds 123, 1, 2 and 3
(MI355X) the averaged profiles give max |u - u_Ghia| 0.021-0.038 (SPH) and
0.015-0.020 (SPH + TVF), max |v - v_Ghia| 0.022-0.035 and 0.015-0.020.  The
bounds sit about 30% above the largest of these.
"""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

js = pytest.importorskip("jax_sph_amd")
from jax_sph_amd import case_setup as cs  # noqa: E402
from jax_sph_amd import validation as V  # noqa: E402
from jax_sph_amd.utils import sph_interpolator  # noqa: E402

REFS = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "validation_refs.json")))
# reverse Poiseuille flow: max |u - u_ref| / max u_ref, measured (MI355X)
# 0.0056 / 0.0047 / 0.0014 / 0.0012 against the series and 0.0144 / 0.0064 /
# 0.0046 / 0.0045 against the reference's profiles at t = 0.2 / 1 / 2 / 10
BOUND_RPF_SERIES = {0.2: 0.0075, 1.0: 0.0065, 2.0: 0.002, 10.0: 0.0017}
BOUND_RPF_DATA = {0.2: 0.019, 1.0: 0.0085, 2.0: 0.006, 10.0: 0.006}


def _at(res, t):
    return int(np.argmin(np.abs(res["t"] - t)))


def test_tgv2d_energy_follows_the_analytic_decay():
    errs = {}
    for nx in (50, 100):
        res = V.tgv_energy(dim=2, nx=nx, Re=100.0, tvf=1.0, t_end=1.0, every=100)
        theory = 0.25 * np.exp(-16 * np.pi ** 2 * res["t"] / 100.0)
        err = res["E"] / theory - 1
        assert abs(err[0]) < 1e-5
        assert np.all(np.diff(res["E"]) < 0), "kinetic energy must decay"
        errs[nx] = (abs(err[_at(res, 0.5)]), abs(err[-1]))
    assert errs[50][0] < 0.06 and errs[50][1] < 0.11, errs
    assert errs[100][0] < 0.035 and errs[100][1] < 0.06, errs
    assert errs[100][1] < errs[50][1]  # converges


def test_tgv3d_re100_energy_against_jax_fluids():
    ref = REFS["tgv3d_re100"]
    rt, rE = np.r_[0.0, ref["t"]], np.r_[0.125, ref["E"]]
    worst = {}
    for nx in (24, 32):
        res = V.tgv_energy(dim=3, nx=nx, Re=100.0, tvf=1.0, t_end=10.0, every=50)
        err = res["E"] / np.interp(res["t"], rt, rE) - 1
        assert abs(res["E"][0] - 0.125) < 1e-5
        assert np.all(np.diff(res["E"]) < 0), "kinetic energy must decay"
        early = res["t"] <= 2.0
        assert np.abs(err[early]).max() < 0.045, (nx, np.abs(err[early]).max())
        worst[nx] = np.abs(err).max()
    assert worst[24] < 0.25 and worst[32] < 0.19, worst
    assert worst[32] < worst[24]  # converges towards the reference solution


@pytest.mark.parametrize("tvf,bu,bv", [(0.0, 0.05, 0.046), (1.0, 0.027, 0.027)])
def test_ldc_re100_centrelines_against_ghia(tvf, bu, bv):
    g = REFS["ldc_re100_ghia"]
    res = V.ldc_centrelines(dx=0.02, t_end=20.0, tvf=tvf, average=5.0)
    assert res["snapshots"] == 6
    u = np.interp(g["y"], res["y"], res["u"])
    v = np.interp(g["x"], res["x"], res["v"])
    assert np.abs(u - np.array(g["u"])).max() < bu
    assert np.abs(v - np.array(g["v"])).max() < bv
    # the lid drives the top of the vertical centreline; the core turns back
    assert res["u"][-1] > 0.6 and res["u"].min() < -0.15


def test_rpf_profiles_against_the_series_and_reference_data():
    """Reverse Poiseuille flow (cases/rpf.py at rpf.yaml's dx = 0.025, from
    the lattice): the row-averaged profile u(y) at t = 0.2, 1, 2 and 10
    against the closed-form start-up series and against the reference's
    profiles (validation/ref/rpf_t_{2,10,20,100}.csv)."""
    refs = REFS["rpf_profiles"]
    res = V.rpf_profiles()
    y = res["y"]
    errs = {}
    for t, prof in sorted(res["u"].items()):
        s = V.rpf_series(y, t, res["g"], res["nu"])
        d = refs[repr(float(t))]
        yd, ud = np.asarray(d["y"]), np.asarray(d["u"])
        sph_at_d = np.interp(yd, np.r_[0.0, y, 1.0], np.r_[0.0, prof, 0.0])
        errs[t] = (np.abs(prof - s).max() / s.max(), np.abs(sph_at_d - ud).max() / ud.max())
    print("rpf max |u - series| / u_max, max |u - data| / u_max:", errs,
          "max |u_y|:", res["uy_max"])
    for t, (e_series, e_data) in errs.items():
        assert e_series < BOUND_RPF_SERIES[t], (t, e_series)
        assert e_data < BOUND_RPF_DATA[t], (t, e_data)
    assert res["uy_max"] < 5e-4                      # measured 3.6e-4


def test_dam_break_front_against_martin_moyce():
    """Surge front of a column of width a = 1 and height 2a collapsing
    (validation/db2d.sh's Riemann settings, dx = 0.02) against Martin &
    Moyce's (1952) experiments (validation/ref/dambreak_data.py): Z = front / a
    at their times T = t sqrt(2 g / a).  Measured (MI355X): within 0.2% for
    T <= 0.62, then trailing the experiment by up to 15.9% near T = 1.8 and
    running parallel to it afterwards (slope 1.46 against 1.44 for T >= 2);
    the same at dx = 0.01 and 0.00625 (within 1%), so the lag is the model's,
    not the resolution's.  The reference compares these by eye only."""
    ref = REFS["db_front_martin_moyce"]
    res = V.db_front()
    Tm, Zm = np.asarray(ref["T"]), np.asarray(ref["Z"])
    z = np.interp(Tm, res["T"], res["Z"])
    rel = (z - Zm) / Zm
    print("dam break front: Z_sph / Z_MM - 1 at T", dict(zip(Tm.tolist(), rel.round(4).tolist())))
    assert res["T"][-1] >= Tm[-1]
    early, late = Tm <= 0.62, Tm >= 2.0
    assert np.abs(rel[early]).max() < 0.01, rel[early]          # measured 0.002
    assert np.abs(rel).max() < 0.2, np.abs(rel).max()            # measured 0.159
    slope_sph = np.polyfit(Tm[late], z[late], 1)[0]
    slope_mm = np.polyfit(Tm[late], Zm[late], 1)[0]
    assert abs(slope_sph / slope_mm - 1) < 0.1, (slope_sph, slope_mm)


def test_interpolator_reproduces_a_smooth_field():
    """utils.py:392-401: Shepard-normalised kernel sums of the initial TGV
    field at off-lattice points recover the analytic field to O(h^2)."""
    cfg = cs.load_config("tgv", {"case": {"dim": 2, "dx": 0.02}})
    (cfg, box, state, *_) = cs.load_case("tgv")(cfg).initialize()
    interp = sph_interpolator(cfg, state, "vector")
    rng = np.random.default_rng(0)
    pts = rng.uniform(0.1, 0.9, size=(64, 2))
    u = interp(state, pts, "u", dim_ind=0).cpu().numpy()
    v = interp(state, pts, "u", dim_ind=1).cpu().numpy()
    x, y = pts[:, 0], pts[:, 1]
    np.testing.assert_allclose(u, -np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y), atol=0.02)
    np.testing.assert_allclose(v, np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y), atol=0.02)
    p = sph_interpolator(cfg, state, "scalar")(state, pts, "rho").cpu().numpy()
    np.testing.assert_allclose(p, 1.0, atol=1e-5)

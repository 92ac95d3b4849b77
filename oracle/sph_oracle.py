"""Oracle: SPH kernels, equations of state, WCSPH right-hand side, si_euler.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, never by the product package.

A numpy restatement, vectorised over the sparse pair list like the reference,
of

* ``jax_sph/kernel.py:26-201``  (all seven kernels; ``grad_w`` in closed form,
  the derivative ``jax.grad`` produces away from the knots)
* ``jax_sph/eos.py:20-57``      (Tait, RIEMANN)
* ``jax_sph/solver.py:24-30,108-256,316-610,705-949`` (``WCSPH.forward``:
  summation / SPH continuity / Riemann continuity, EOS, GWBC, Riemann wall BC,
  heat conduction, standard and Riemann accelerations, transport velocity,
  artificial viscosity)
* ``jax_sph/integrator.py:8-58`` (``si_euler``)

Pair convention follows ``solver.py:722``: ``i_s, j_s = idx``; every
``segment_sum`` accumulates onto ``i_s``; ``dr_ij = r[i_s] - r[j_s]`` (minimum
image).  Padded pairs (index N) are dropped, as the reference's segment_sum does.

Scalar constants are rounded the way the reference's weak-typed JAX scalars
are: a Python float meeting a float32 array is rounded to float32 first.

How each part is pinned (tests/test_oracle_physics.py):
* kernels: closed-form integrals (2D/3D normalisation, incl. the reference's
  3/(359 pi) = 1.0027855 quirk), the reference's 1D half-integral / sign tests
  (``tests/test_kernel.py``), central finite differences of ``w`` vs ``grad_w``;
* EOS: ``rho_fn(p_fn(rho)) == rho``, p(rho_ref) == p_bg, closed forms;
* density summation: lattice sum vs the continuum integral of W;
* forces: momentum conservation sum(m a) = 0 (periodic, no walls) for the
  standard, transport-velocity, Riemann and artificial-viscosity terms, energy
  conservation sum(m Cp dT/dt) = 0 for heat, zero force on a uniform lattice at
  rest, Galilean invariance of the viscous terms;
* GWBC / Riemann wall BC: exact Shepard identities (uniform fluid velocity U
  gives wall velocity 2*u_w - U; uniform p gives wall p), hydrostatic
  extrapolation with gravity;
* si_euler: free streaming with zero RHS, the kick/drift ordering on a
  hand-computed step;
* end to end: 2D Taylor-Green kinetic-energy decay vs exp(-16 pi^2 t / Re)
  (``validation/validate.py:181-194``).
"""
from __future__ import annotations

import math
from typing import Callable, Dict, Optional

import numpy as np

from .nl_oracle import distance, periodic_displacement, periodic_shift

EPS = float(np.finfo(float).eps)  # solver.py:21

FLUID, SOLID_WALL, MOVING_WALL, DIRICHLET_WALL = 0, 1, 2, 3
WALL_TAGS = (SOLID_WALL, MOVING_WALL, DIRICHLET_WALL)  # utils.py:32


def _f(t, x):
    return t(x)


# ------------------------------------------------------------------------ kernels
class Kernel:
    """One SPH kernel: ``w(r)``, ``grad_w(r)`` and ``cutoff`` (kernel.py)."""

    def __init__(self, name: str, h: float, dim: int, dtype=np.float32):
        self.name, self.dim = name, dim
        self.t = np.dtype(dtype).type
        self.one_over_h = 1.0 / h
        if name == "CSK":                                          # kernel.py:26-48
            nc = 2.0
            sig = {1: 2.0 / 3.0 * self.one_over_h,
                   2: 10.0 / 7.0 / math.pi * self.one_over_h ** 2,
                   3: 1.0 / math.pi * self.one_over_h ** 3}[dim]
        elif name == "QSK":                                        # kernel.py:51-72
            nc = 3.0
            sig = {1: 1.0 / 120.0 * self.one_over_h,
                   2: 7.0 / 478.0 / math.pi * self.one_over_h ** 2,
                   3: 3.0 / 359.0 / math.pi * self.one_over_h ** 3}[dim]
        elif name == "WC2K":                                       # kernel.py:75-103
            nc = 2.0
            sig = {1: 5.0 / 8.0 * self.one_over_h,
                   2: 7.0 / 4.0 / math.pi * self.one_over_h ** 2,
                   3: 21.0 / 16.0 / math.pi * self.one_over_h ** 3}[dim]
        elif name == "WC4K":                                       # kernel.py:106-134
            nc = 2.0
            sig = {1: 3.0 / 4.0 * self.one_over_h,
                   2: 9.0 / 4.0 / math.pi * self.one_over_h ** 2,
                   3: 495.0 / 256.0 / math.pi * self.one_over_h ** 3}[dim]
        elif name == "WC6K":                                       # kernel.py:137-165
            nc = 2.0
            sig = {1: 55.0 / 64.0 * self.one_over_h,
                   2: 78.0 / 28.0 / math.pi * self.one_over_h ** 2,
                   3: 1365.0 / 512.0 / math.pi * self.one_over_h ** 3}[dim]
        elif name in ("GK", "SGK"):                                # kernel.py:168-201
            nc = 3.0
            sig = 1.0 / math.pi ** (dim / 2) * self.one_over_h ** dim
        else:
            raise ValueError(name)
        self.cutoff = nc * h
        self.sigma = sig

    def _q(self, r):
        t = self.t
        return r * t(self.one_over_h)

    def w(self, r):
        t, q = self.t, self._q(np.asarray(r, dtype=self.t))
        s = t(self.sigma)
        one, zero = t(1.0), t(0.0)
        if self.name == "QSK":
            q1 = np.maximum(zero, one - q)
            q2 = np.maximum(zero, t(2.0) - q)
            q3 = np.maximum(zero, t(3.0) - q)
            return s * (q3 ** 5 - t(6.0) * q2 ** 5 + t(15.0) * q1 ** 5)
        if self.name == "CSK":
            c1 = (one - q >= 0).astype(self.t)
            c2 = ((t(2) - q < 1) & (t(2) - q >= 0)).astype(self.t)
            q1 = one - t(1.5) * q ** 2 * (one - q / t(2))
            q2 = t(0.25) * (t(2) - q) ** 3
            return s * (q1 * c1 + q2 * c2)
        if self.name in ("WC2K", "WC4K", "WC6K"):
            q1 = np.maximum(zero, one - t(0.5) * q)
            if self.dim == 1:
                if self.name == "WC2K":
                    return s * (q1 ** 3 * (t(1.5) * q + one))
                if self.name == "WC4K":
                    return s * (q1 ** 5 * (t(2.0) * q ** 2 + t(2.5) * q + one))
                return s * (q1 ** 7 * (t(21.0 / 8.0) * q ** 3 + t(19.0 / 4.0) * q ** 2
                                       + t(3.5) * q + one))
            if self.name == "WC2K":
                return s * (q1 ** 4 * (t(2.0) * q + one))
            if self.name == "WC4K":
                return s * (q1 ** 6 * (t(35.0 / 12.0) * q ** 2 + t(3) * q + one))
            return s * (q1 ** 8 * (t(4.0) * q ** 3 + t(6.25) * q ** 2 + t(4) * q + one))
        if self.name == "GK":
            c = (t(3) - q >= 0).astype(self.t)
            return s * c * np.exp(-(q ** 2))
        if self.name == "SGK":
            c = (t(3) - q >= 0).astype(self.t)
            return s * c * np.exp(-(q ** 2)) * (t(self.dim / 2 + 1) - q ** 2)
        raise ValueError(self.name)

    def grad_w(self, r):
        """dW/dr in closed form (equals jax.grad(w) wherever W is differentiable)."""
        t, q = self.t, self._q(np.asarray(r, dtype=self.t))
        s, ih = t(self.sigma), t(self.one_over_h)
        one, zero = t(1.0), t(0.0)
        if self.name == "QSK":
            q1 = np.maximum(zero, one - q)
            q2 = np.maximum(zero, t(2.0) - q)
            q3 = np.maximum(zero, t(3.0) - q)
            return -(t(5.0) * s * ih) * (q3 ** 4 - t(6.0) * q2 ** 4 + t(15.0) * q1 ** 4)
        if self.name == "CSK":
            c1 = (one - q >= 0).astype(self.t)
            c2 = ((t(2) - q < 1) & (t(2) - q >= 0)).astype(self.t)
            d1 = -t(3.0) * q + t(2.25) * q ** 2
            d2 = -t(0.75) * (t(2) - q) ** 2
            return s * ih * (d1 * c1 + d2 * c2)
        if self.name in ("WC2K", "WC4K", "WC6K"):
            q1 = np.maximum(zero, one - t(0.5) * q)
            if self.dim == 1:
                if self.name == "WC2K":   # d/dq[(1-q/2)^3 (1.5q+1)] = -3 q (1-q/2)^2
                    return s * ih * (-t(3.0) * q * q1 ** 2)
                if self.name == "WC4K":
                    p = t(2.0) * q ** 2 + t(2.5) * q + one
                    dp = t(4.0) * q + t(2.5)
                    return s * ih * (-t(2.5) * q1 ** 4 * p + q1 ** 5 * dp)
                p = t(21.0 / 8.0) * q ** 3 + t(19.0 / 4.0) * q ** 2 + t(3.5) * q + one
                dp = t(63.0 / 8.0) * q ** 2 + t(9.5) * q + t(3.5)
                return s * ih * (-t(3.5) * q1 ** 6 * p + q1 ** 7 * dp)
            if self.name == "WC2K":       # d/dq[(1-q/2)^4 (2q+1)] = -5 q (1-q/2)^3
                return -(t(5.0) * s * ih) * q * q1 ** 3
            if self.name == "WC4K":
                p = t(35.0 / 12.0) * q ** 2 + t(3) * q + one
                dp = t(35.0 / 6.0) * q + t(3)
                return s * ih * (-t(3.0) * q1 ** 5 * p + q1 ** 6 * dp)
            p = t(4.0) * q ** 3 + t(6.25) * q ** 2 + t(4) * q + one
            dp = t(12.0) * q ** 2 + t(12.5) * q + t(4)
            return s * ih * (-t(4.0) * q1 ** 7 * p + q1 ** 8 * dp)
        if self.name == "GK":
            c = (t(3) - q >= 0).astype(self.t)
            return s * ih * c * (-t(2) * q) * np.exp(-(q ** 2))
        if self.name == "SGK":
            c = (t(3) - q >= 0).astype(self.t)
            e = np.exp(-(q ** 2))
            a = t(self.dim / 2 + 1)
            return s * ih * c * e * (-t(2) * q * (a - q ** 2) - t(2) * q)
        raise ValueError(self.name)


# --------------------------------------------------------------------------- EOS
class TaitEoS:
    """``eos.py:20-38``."""

    def __init__(self, p_ref, rho_ref, p_background, gamma, dtype=np.float32):
        self.t = np.dtype(dtype).type
        self.p_ref, self.rho_ref, self.p_bg, self.gamma = p_ref, rho_ref, p_background, gamma

    def p_fn(self, rho):
        t = self.t
        x = rho / t(self.rho_ref)
        x = x if self.gamma == 1.0 else x ** t(self.gamma)
        return t(self.p_ref) * (x - t(1)) + t(self.p_bg)

    def rho_fn(self, p):
        t = self.t
        p_temp = p + t(self.p_ref) - t(self.p_bg)
        x = p_temp / t(self.p_ref)
        x = x if self.gamma == 1.0 else x ** t(1 / self.gamma)
        return t(self.rho_ref) * x


class RIEMANNEoS:
    """``eos.py:41-57``."""

    def __init__(self, rho_ref, p_background, u_ref, dtype=np.float32):
        self.t = np.dtype(dtype).type
        self.rho_ref, self.p_bg, self.u_ref = rho_ref, p_background, u_ref

    def p_fn(self, rho):
        t = self.t
        return t(100 * self.u_ref ** 2) * (rho - t(self.rho_ref)) + t(self.p_bg)

    def rho_fn(self, p):
        t = self.t
        return (p - t(self.p_bg)) / t(100 * self.u_ref ** 2) + t(self.rho_ref)


# -------------------------------------------------------------------- segment sum
def _seg(vals, seg, N):
    out = np.zeros((N,) + vals.shape[1:], dtype=vals.dtype)
    np.add.at(out, seg, vals)
    return out


def _isin_wall(tag):
    return np.isin(tag, WALL_TAGS)


# ------------------------------------------------------------------------- WCSPH
class WCSPH:
    """``solver.py:613-951`` restated (SPH, RIE and DELTA solvers, density
    renormalisation)."""

    def __init__(self, box, eos, g_ext_fn: Callable, dx, dim, dt, c_ref, eta_limiter=3,
                 solver="SPH", kernel="QSK", h_fac=1.0, is_bc_trick=False,
                 is_rho_evol=False, artificial_alpha=0.0, is_free_slip=False,
                 is_heat_conduction=False, dtype=np.float32, diff_delta=0.02, diff_alpha=0.1,
                 is_rho_renorm=False):
        if solver not in ("SPH", "RIE", "DELTA"):
            raise NotImplementedError(f"unknown solver {solver!r}")
        self.diff_delta, self.diff_alpha = diff_delta, diff_alpha      # solver.py:626-627
        self.is_rho_renorm = is_rho_renorm
        self.support = h_fac * dx                                      # solver.py:688, 693
        self.t = np.dtype(dtype).type
        self.dtype = np.dtype(dtype)
        self.box = np.asarray(box, dtype=self.dtype).reshape(-1)
        self.eos, self.g_ext_fn = eos, g_ext_fn
        self.dx, self.dim, self.dt, self.c_ref = dx, dim, dt, c_ref
        self.eta_limiter = eta_limiter
        self.solver, self.is_bc_trick, self.is_rho_evol = solver, is_bc_trick, is_rho_evol
        self.artificial_alpha = artificial_alpha
        self.is_free_slip, self.is_heat_conduction = is_free_slip, is_heat_conduction
        self.kernel = Kernel(kernel, h_fac * dx, dim, dtype)

    # beta_fn of limiter_fn_wrapper (solver.py:574-588)
    def _beta(self, u_L, u_R):
        t = self.t
        if self.eta_limiter == -1:
            return np.full_like(u_L, t(self.c_ref))
        temp = t(self.eta_limiter) * np.maximum(u_L - u_R, t(0))
        return np.minimum(temp, t(self.c_ref))

    def forward(self, state: Dict[str, np.ndarray], idx: np.ndarray) -> Dict[str, np.ndarray]:
        t = self.t
        eps = t(EPS)
        dt = t(self.dt)
        r, tag, mass, eta = state["r"], state["tag"], state["mass"], state["eta"]
        u, v, dudt, dvdt = state["u"], state["v"], state["dudt"], state["dvdt"]
        rho, drhodt, p = state["rho"], state["drhodt"], state["p"]
        nw, kappa, Cp = state["nw"], state["kappa"], state["Cp"]
        T, dTdt = state["T"], state["dTdt"]
        N = len(r)

        idx = np.asarray(idx)
        keep = idx[0] < N
        i_s, j_s = idx[0][keep].astype(np.int64), idx[1][keep].astype(np.int64)

        dr = periodic_displacement(self.box, r[i_s] - r[j_s])          # solver.py:724
        dist = distance(dr)
        w_dist = self.kernel.w(dist)
        e_s = dr / (dist[:, None] + eps)
        gw = self.kernel.grad_w(dist)
        grad_w_dist = gw[:, None] * e_s
        g_ext = self.g_ext_fn(r)

        wall_mask = np.where(_isin_wall(tag), t(1.0), t(0.0))
        fluid_mask = np.where(tag == FLUID, t(1.0), t(0.0))
        wall_j = _isin_wall(tag)[j_s]

        # Riemann velocity BCs (solver.py:741-744, 547-552)
        if self.is_bc_trick and self.solver == "RIE" and not self.is_free_slip:
            wdf = w_dist * fluid_mask[j_s]
            nom = _seg(wdf[:, None] * u[j_s], i_s, N)
            den = _seg(wdf, i_s, N)
            u_tilde = nom / (den[:, None] + eps)
        else:
            u_tilde = u

        # density (solver.py:750-796)
        if self.is_rho_evol and self.solver == "SPH":
            v_j = (mass / rho)[j_s]
            temp = v_j * ((u[i_s] - u[j_s]) * grad_w_dist).sum(axis=1)
            drhodt = rho * _seg(temp, i_s, N)
            rho = rho + dt * drhodt
            if self.is_rho_renorm:
                rho = self._rho_renorm(rho, mass, i_s, j_s, w_dist, N)
        elif self.is_rho_evol and self.solver == "DELTA":
            rho, drhodt = self._rho_evol_delta(rho, mass, dr, dist, u, i_s, j_s, dt, N,
                                               fluid_mask[j_s])
            if self.is_rho_renorm:
                rho = self._rho_renorm(rho, mass, i_s, j_s, w_dist, N)
        elif self.is_rho_evol and self.solver == "RIE":
            temp = self._rho_evol_riemann(e_s, rho, mass, u, p, dr, gw, wall_j, nw,
                                          g_ext, i_s, j_s)
            drhodt = _seg(temp, i_s, N) * fluid_mask
            rho = rho + dt * drhodt
            if self.is_rho_renorm:
                rho = self._rho_renorm(rho, mass, i_s, j_s, w_dist, N)
        else:
            rho_ = mass * _seg(w_dist, i_s, N)
            rho = np.where(fluid_mask != 0, rho_, rho)

        p = self.eos.p_fn(rho)                                           # solver.py:801
        p_bg_tvf = self.eos.p_fn(np.zeros_like(p))

        if self.is_bc_trick and self.solver in ("SPH", "DELTA"):           # solver.py:806
            p, rho, u, v, T = self._gwbc(T, rho, tag, u, v, p, g_ext, i_s, j_s, w_dist,
                                         dr, nw, N)
            mask = None
        elif self.is_bc_trick and self.solver == "RIE":
            mask = fluid_mask[i_s] if self.is_free_slip else np.ones(len(i_s), self.dtype)
            if self.is_heat_conduction:
                T = self._heat_bc(fluid_mask[j_s], w_dist, T, i_s, j_s, tag, N)
        else:
            mask = np.ones(len(i_s), self.dtype)

        if self.is_heat_conduction:                                      # solver.py:832-850
            T = T + dt * dTdt
            keff = (kappa[i_s] * kappa[j_s]) / (kappa[i_s] + kappa[j_s])
            F = (dr * (gw[:, None] * e_s)).sum(axis=1) / (dist * dist + eps)
            out = (t(4) * mass[j_s] * keff * (T[i_s] - T[j_s]) * F) / (
                Cp[i_s] * rho[i_s] * rho[j_s])
            dTdt = _seg(out, i_s, N)

        # common prefactor (solver.py:246-248)
        m_i, m_j = mass[i_s], mass[j_s]
        rho_i, rho_j = rho[i_s], rho[j_s]
        wv = ((m_i / rho_i) ** 2 + (m_j / rho_j) ** 2) / m_i
        c = wv * gw / (dist + eps)

        if self.solver in ("SPH", "DELTA"):                              # solver.py:221-256
            eta_ij = t(2) * eta[i_s] * eta[j_s] / (eta[i_s] + eta[j_s] + eps)
            p_ij = (rho_j * p[i_s] + rho_i * p[j_s]) / (rho_i + rho_j)
            Ar_i = (rho_i[:, None] * u[i_s]) * ((v[i_s] - u[i_s]) * dr).sum(axis=1)[:, None]
            Ar_j = (rho_j[:, None] * u[j_s]) * ((v[j_s] - u[j_s]) * dr).sum(axis=1)[:, None]
            A_r = (Ar_i + Ar_j) / t(2)
            a = c[:, None] * (-p_ij[:, None] * dr + A_r + eta_ij[:, None] * (u[i_s] - u[j_s]))
            if self.solver == "DELTA":                                   # solver.py:290-311
                kg = gw[:, None] * (dr / (dist[:, None] + eps))
                V_j = m_j / rho_j
                pi_ij = ((u[j_s] - u[i_s]) * -dr).sum(axis=1) / (dist + eps) ** 2
                ad = (V_j * pi_ij)[:, None] * kg
                ad = ad * t(self.diff_alpha)
                ad = ad * t(self.support)
                ad = ad * t(self.c_ref)
                ad = ad * t(self.eos.rho_ref)
                ad = ad / rho_i[:, None]
                ad = ad * fluid_mask[j_s][:, None]
                a = a + ad
        else:
            a = self._acc_riemann(e_s, dr, dist, gw, rho, mass, u, p, eta, wall_j, mask,
                                  nw, g_ext, u_tilde, i_s, j_s)
        dudt = _seg(a, i_s, N)

        out_tv = (c * t(1.0) * p_bg_tvf[i_s])[:, None] * dr             # solver.py:199-213
        dvdt = _seg(out_tv, i_s, N)

        if self.artificial_alpha != 0.0:                                 # solver.py:404-428
            h_ab = self.dx
            c_ab = 10.0
            rho_ab = (rho_i + rho_j) / t(2)
            num = m_j * t(self.artificial_alpha) * t(h_ab) * t(c_ab)
            num = num * ((u[i_s] - u[j_s]) * dr).sum(axis=1)
            num = num[:, None] * grad_w_dist
            den = (rho_ab * (dist ** 2 + t(0.01 * h_ab ** 2)))[:, None]
            mf = tag == FLUID
            mfe = (mf[j_s] & mf[i_s]).astype(self.dtype)
            dudt = dudt + _seg(mfe[:, None] * num / den, i_s, N)

        return {
            "r": r, "tag": tag, "u": u, "v": v, "drhodt": drhodt,
            "dudt": dudt + g_ext, "dvdt": dvdt, "rho": rho, "p": p, "mass": mass,
            "eta": eta, "dTdt": dTdt, "T": T, "kappa": kappa, "Cp": Cp, "nw": nw,
        }

    # solver.py:167-173 (all particles, no fluid mask)
    def _rho_renorm(self, rho, mass, i_s, j_s, w_dist, N):
        t = self.t
        nominator = _seg(mass[j_s] * w_dist, i_s, N)
        den = _seg((mass / rho)[j_s] * w_dist, i_s, N)
        den = np.where(den > 1, t(1), den)
        return nominator / den

    # solver.py:33-105, literally: L_matj is the inverse of the j_s-grouped sum
    # of outer(r_ij, grad W_ij V_i), and rho_grad_term_j is segment-summed over
    # i_s, then read at j_s.
    def _rho_evol_delta(self, rho, mass, r_ij, d_ij, u, i_s, j_s, dt, N, fluidmask_j):
        t = self.t
        eps = t(EPS)
        kg, V_i, V_j, L_i, L_j, rgi, rgj = self.delta_gradients(rho, mass, r_ij, d_ij, i_s,
                                                                 j_s, N)
        rho_term = t(2) * (rho[j_s] - rho[i_s])[:, None] * (-r_ij) / (
            (d_ij + eps) ** 2)[:, None]
        psi = rho_term - rgi[i_s] - rgj[j_s]
        diff_term = _seg((psi * kg).sum(axis=1) * V_j * fluidmask_j, i_s, N)
        cont = ((u[i_s] - u[j_s]) * kg).sum(axis=1) * V_j
        coef = t(self.c_ref * self.diff_delta * self.support)
        drhodt = rho * _seg(cont, i_s, N) + coef * diff_term
        return rho + t(dt) * drhodt, drhodt

    def delta_gradients(self, rho, mass, r_ij, d_ij, i_s, j_s, N):
        """Renormalisation matrices and density-gradient terms of solver.py:42-69."""
        t = self.t
        eps = t(EPS)
        dim = r_ij.shape[1]
        e_ab = r_ij / (d_ij[:, None] + eps)
        kg = self.kernel.grad_w(d_ij)[:, None] * e_ab
        V_i = mass[i_s] / rho[i_s]
        V_j = mass[j_s] / rho[j_s]

        def seg_mat(vals, seg):
            out = np.zeros((N, dim, dim), self.dtype)
            np.add.at(out, seg, vals)
            return out

        temp = (-r_ij)[:, :, None] * (kg * V_j[:, None])[:, None, :]   # L_fn(r_ij, kg, V_j)
        L_i = np.linalg.inv(seg_mat(temp, i_s)).astype(self.dtype)
        temp = (r_ij)[:, :, None] * (kg * V_i[:, None])[:, None, :]    # L_fn(-r_ij, kg, V_i)
        L_j = np.linalg.inv(seg_mat(temp, j_s)).astype(self.dtype)
        gi = (rho[j_s] - rho[i_s])[:, None] * np.einsum("pab,pb->pa", L_i[i_s],
                                                          kg * V_j[:, None])
        rgi = _seg(gi, i_s, N)
        gj = (rho[i_s] - rho[j_s])[:, None] * np.einsum("pab,pb->pa", L_j[j_s],
                                                          -kg * V_i[:, None])
        rgj = _seg(gj, i_s, N)
        return kg, V_i, V_j, L_i, L_j, rgi, rgj

    # solver.py:147-152 and 359-364: the right state of a wall neighbour is the
    # left (fluid) state extrapolated hydrostatically to the wall particle,
    # p_R = p_L + rho_L g_i.(-r_ij), rho_R = rho_fn(p_R); a fluid neighbour's is
    # its own state
    def _riemann_right_state(self, wall_j, p_L, rho_L, g_i, dr, p_j, rho_j):
        p_R = np.where(wall_j, p_L + rho_L * (g_i * -dr).sum(1), p_j)
        rho_R = np.where(wall_j, self.eos.rho_fn(p_R), rho_j)
        return p_R, rho_R

    # solver.py:108-164
    def _rho_evol_riemann(self, e_s, rho, mass, u, p, dr, gw, wall_j, nw, g_ext, i_s, j_s):
        t = self.t
        kg = gw[:, None] * e_s
        u_i, u_j = u[i_s], u[j_s]
        n_j = nw[j_s]
        u_L = np.where(wall_j, (u_i * -n_j).sum(1), (u_i * -e_s).sum(1))
        p_L, rho_L = p[i_s], rho[i_s]
        u_R = np.where(wall_j, -u_L + t(2) * (u_j * n_j).sum(1), (u_j * -e_s).sum(1))
        p_R, rho_R = self._riemann_right_state(wall_j, p_L, rho_L, g_ext[i_s], dr, p[j_s],
                                               rho[j_s])
        U_avg = (u_L + u_R) / t(2)
        v_avg = (u_i + u_j) / t(2)
        rho_avg = (rho_L + rho_R) / t(2)
        U_star = U_avg + t(0.5) * (p_L - p_R) / (rho_avg * t(self.c_ref))
        v_star = U_star[:, None] * -e_s + (v_avg - U_avg[:, None] * -e_s)
        return t(2) * rho[i_s] * mass[j_s] / rho[j_s] * ((u_i - v_star) * kg).sum(1)

    # solver.py:316-401 (the A term is identically zero there: stress(rho,u,u)=0)
    def _acc_riemann(self, e_s, dr, dist, gw, rho, mass, u, p, eta, wall_j, mask, nw,
                     g_ext, u_tilde, i_s, j_s):
        t = self.t
        eps = t(EPS)
        kg = gw[:, None] * e_s
        u_i, u_j = u[i_s], u[j_s]
        n_j = nw[j_s]
        u_L = np.where(wall_j, (u_i * -n_j).sum(1), (u_i * -e_s).sum(1))
        p_L, rho_L = p[i_s], rho[i_s]
        u_R = np.where(wall_j, -u_L + t(2) * (u_j * n_j).sum(1), (u_j * -e_s).sum(1))
        p_R, rho_R = self._riemann_right_state(wall_j, p_L, rho_L, g_ext[i_s], dr, p[j_s],
                                               rho[j_s])
        P_avg = (p_L + p_R) / t(2)
        rho_avg = (rho_L + rho_R) / t(2)
        eta_ij = t(2) * eta[i_s] * eta[j_s] / (eta[i_s] + eta[j_s] + eps)
        P_star = P_avg + t(0.5) * rho_avg * (u_L - u_R) * self._beta(u_L, u_R)
        rho_i, rho_j, m_j = rho[i_s], rho[j_s], mass[j_s]
        eq_9 = (-t(2) * m_j * (P_star / (rho_i * rho_j)))[:, None] * kg
        u_d = t(2) * u_j - u_tilde[j_s]
        v_ij = np.where(wall_j[:, None], u_i - u_d, u_i - u_j)
        eq_6 = (t(2) * m_j * eta_ij / (rho_i * rho_j))[:, None] * v_ij / (dist + eps)[:, None]
        eq_6 = eq_6 * (gw * mask)[:, None]
        return eq_9 + eq_6

    # solver.py:431-530
    def _gwbc(self, T, rho, tag, u, v, p, g_ext, i_s, j_s, w_dist, dr, nw, N):
        t = self.t
        eps = t(EPS)
        mask_bc = _isin_wall(tag)
        mjf = np.where(tag[j_s] == FLUID, t(1.0), t(0.0))
        wf = w_dist * mjf
        wsum = _seg(wf, i_s, N)

        def shepard(x):
            return _seg(wf[:, None] * x[j_s], i_s, N) / (wsum[:, None] + eps)

        if self.is_free_slip:
            n_in = -nw

            def bc(x):
                xw = shepard(x)
                xw = n_in * (xw * n_in).sum(axis=1, keepdims=True)
                return np.where(mask_bc[:, None], t(2) * x - xw, x)
        else:
            def bc(x):
                return np.where(mask_bc[:, None], t(2) * x - shepard(x), x)

        u = bc(u)
        v = bc(v)
        p_un = _seg(wf * p[j_s], i_s, N)
        rwf = _seg((rho[j_s] * wf)[:, None] * dr, i_s, N)
        p_ext = (g_ext * rwf).sum(axis=1)
        p_wall = (p_un + p_ext) / (wsum + eps)
        p = np.where(mask_bc, p_wall, p)
        rho = self.eos.rho_fn(p)
        if self.is_heat_conduction:
            tw = _seg(wf * T[j_s], i_s, N) / (wsum + eps)
            m = np.isin(tag, (SOLID_WALL, MOVING_WALL))
            T = np.where(m, tw, T)
        return p, rho, u, v, T

    # solver.py:554-565
    def _heat_bc(self, mjf, w_dist, T, i_s, j_s, tag, N):
        t = self.t
        wf = w_dist * mjf
        wsum = _seg(wf, i_s, N)
        tw = _seg(wf * T[j_s], i_s, N) / (wsum + t(EPS))
        m = np.isin(tag, (SOLID_WALL, MOVING_WALL))
        return np.where(m, tw, T)


# --------------------------------------------------------------------- integrator
def si_euler(tvf, model, box, bc_fn: Optional[Callable], nl_update, dtype=np.float32,
             nw_fn: Optional[Callable] = None):
    """``integrator.py:8-58``: returns advance(dt, state, nbrs) -> (state, nbrs)."""
    t = np.dtype(dtype).type
    box = np.asarray(box, dtype=dtype).reshape(-1)

    def advance(dt, state, nbrs):
        state = dict(state)
        dtf = t(dt)
        state["u"] = state["u"] + t(1.0 * dtf) * state["dudt"]
        state["v"] = state["u"] + t(t(tvf * 0.5) * dtf) * state["dvdt"]
        state["r"] = periodic_shift(box, state["r"], t(1.0 * dtf) * state["v"])
        if nw_fn is not None:                                   # integrator.py:32-33
            state["nw"] = nw_fn(state["r"])
        nbrs = nl_update(state["r"], nbrs)
        state = model(state, nbrs.idx)
        if bc_fn is not None:
            state = bc_fn(state)
        return state, nbrs

    return advance


def wall_normals(r, tag, layer, offset_vec, box, dtype=np.float32):
    """``utils.py:197-275``: normal of each wall particle = unit periodic
    displacement (layer - wall) to its nearest point of the fine wall layer
    (brute force over the layer; first minimum on ties)."""
    t = np.dtype(dtype).type
    walls = np.isin(tag, (1, 2, 3))
    rw = (r[walls] - np.asarray(offset_vec, dtype)).astype(dtype)
    lay = np.asarray(layer, dtype)
    d = periodic_displacement(np.asarray(box, dtype), lay[None, :, :] - rw[:, None, :])
    dist = np.sqrt((d * d).sum(-1))
    k = np.argmin(dist, axis=1)
    rows = np.arange(len(k))
    nw = np.zeros_like(r, dtype=dtype)
    nw[walls] = d[rows, k] / (dist[rows, k] + t(np.finfo(float).eps))[:, None]
    return nw


def get_ekin(state, dx):
    """``utils.py:128-133``: 0.5 * sum(v_fluid^2) * dx^d (float64 accumulation)."""
    v = state["v"].astype(np.float64)
    vw = np.where((state["tag"] == FLUID)[:, None], v, 0.0)
    return 0.5 * float(np.square(vw).sum()) * dx ** v.shape[1]

"""Pins the BDF / DIRK drivers (hdg_amd/timestep.py; kernel/hdg_solve_bdf.m,
kernel/hdg_solve_dirk.m) on the CPU:

* the BDF coefficients satisfy the order conditions of BDF1-4, and every DIRK
  tableau its row-sum (stage times) and order conditions up to its order;
* the drivers' stage arithmetic (SDG, fc_u, stage times, the final
  combination, the history shift) run with a scalar stand-in for the Newton
  solve -- d_ii U - SH = lam U + g(time), solved exactly -- gives the method's
  stability function R(lam dt) for g = 0, and reproduces polynomial solutions
  of degree <= the order exactly for lam = 0.
"""
import types

import numpy as np
import pytest

from hdg_amd.timestep import bdfcoeff, dirk_tableau, dirkcoeff, hdg_solve_bdf, hdg_solve_dirk

DIRKS = [(1, 1), (1, 2), (2, 2), (2, 3), (3, 3), (3, 4), (5, 5)]


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_bdf_order_conditions(p):
    a = bdfcoeff(p)
    tau = -np.arange(p + 1, dtype=float)          # nodes t_{n+1-i} - t_{n+1} in units of dt
    for k in range(p + 1):
        want = 1.0 if k == 1 else 0.0
        assert abs(np.sum(a * tau ** k) - want) < 1e-13, k
    # one order more fails (the order is exactly p)
    assert abs(np.sum(a * tau ** (p + 1))) > 1e-3


def _order_conditions(a, b, t):
    """Residuals of the Runge-Kutta order conditions, by order."""
    e = np.ones_like(b)
    at, at2 = a @ t, a @ (t ** 2)
    return {1: [b @ e - 1],
            2: [b @ t - 1 / 2],
            3: [b @ t ** 2 - 1 / 3, b @ at - 1 / 6],
            4: [b @ t ** 3 - 1 / 4, b @ (t * at) - 1 / 8, b @ at2 - 1 / 12, b @ (a @ at) - 1 / 24],
            5: [b @ t ** 4 - 1 / 5, b @ (t ** 2 * at) - 1 / 10, b @ (t * at2) - 1 / 15, b @ (t * (a @ at)) - 1 / 30,
                b @ (at * at) - 1 / 20, b @ (a @ t ** 3) - 1 / 20, b @ (a @ (t * at)) - 1 / 40,
                b @ (a @ at2) - 1 / 60, b @ (a @ (a @ at)) - 1 / 120]}


@pytest.mark.parametrize("q,p", DIRKS)
def test_dirk_tableau(q, p):
    a, b, t = dirk_tableau(q, p)
    assert np.allclose(np.triu(a, 1), 0.0)          # diagonally implicit
    tol = 1e-9 if (q, p) == (3, 3) else 1e-12        # a1 = 0.4358665215 is given to 10 digits
    assert np.abs(a.sum(1) - t).max() < tol
    res = _order_conditions(a, b, t)
    for k in range(1, p + 1):
        assert np.abs(res[k]).max() < tol, (k, res[k])
    if p < 5:
        assert np.abs(res[p + 1]).max() > 1e-6      # and not one order more
    d, c, tt = dirkcoeff(q, p)
    assert np.allclose(d @ a, np.eye(q)) and np.allclose(c, b @ np.linalg.inv(a))


class ScalarState:
    """Stand-in for HdgState: every entry of UDG is an independent scalar
    problem fc_u U - SH = lam U + g(time), solved exactly by solve(); the
    trace UH takes the stage's UDG value (so the step must combine both the
    same way)."""

    def __init__(self, u0, lam, g=None):
        self.UDG = types.SimpleNamespace(t=u0.copy())
        self.UH = types.SimpleNamespace(t=u0.copy())
        self.lam = lam
        self.g = g or (lambda time: 0.0)
        self.app = types.SimpleNamespace(time=0.0, fc_u=0.0, fc_q=1.0, tdep=0)
        self.SH = None
        self.calls = []

    def set_source(self, SH):
        self.SH = SH.copy()

    def getq(self):
        pass

    def solve(self, **kw):
        fc, g = self.app.fc_u, self.g(self.app.time)
        self.calls.append((self.app.time, fc))
        self.UDG.t[...] = (self.SH + g) / (fc - self.lam)
        self.UH.t[...] = self.UDG.t
        return None


def _stability(a, b, z):
    q = len(b)
    return 1 + z * b @ np.linalg.solve(np.eye(q) - z * a, np.ones(q))


@pytest.mark.parametrize("q,p", DIRKS)
def test_dirk_driver_stability(q, p):
    dt = 0.3
    lam = -np.array([0.0, 0.5, 1.0, 3.0, 10.0, 100.0])
    u0 = np.linspace(1.0, 2.0, lam.size)
    S = ScalarState(u0, lam)
    hdg_solve_dirk(S, 0.0, dt, q, p)
    a, b, _ = dirk_tableau(q, p)
    want = np.array([_stability(a, b, z) for z in lam * dt]) * u0
    assert np.abs(S.UDG.t - want).max() < 1e-12 * np.abs(want).max()
    assert np.abs(S.UH.t - want).max() < 1e-12 * np.abs(want).max()
    assert [c[1] for c in S.calls] == pytest.approx(list(np.diag(dirkcoeff(q, p)[0]) / dt))


@pytest.mark.parametrize("q,p", DIRKS)
def test_dirk_driver_polynomial_exact(q, p):
    """u' = P'(t) with P of degree p: the step lands on P(t + dt) exactly,
    which needs the stage times time + t_i dt."""
    rng = np.random.default_rng(q * 10 + p)
    coef = rng.standard_normal(p + 1)
    P = np.polynomial.Polynomial(coef)
    dP = P.deriv()
    t0, dt = 0.7, 0.25
    S = ScalarState(np.array([P(t0)]), np.zeros(1), g=lambda tt: dP(tt))
    hdg_solve_dirk(S, t0, dt, q, p)
    tol = 1e-8 if (q, p) == (3, 3) else 1e-12
    assert abs(S.UDG.t[0] - P(t0 + dt)) < tol * max(1.0, abs(P(t0 + dt)))


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_bdf_driver_polynomial_exact(p):
    rng = np.random.default_rng(p)
    P = np.polynomial.Polynomial(rng.standard_normal(p + 1))
    dP = P.deriv()
    t0, dt = 0.4, 0.1
    # history U(t0), U(t0 - dt), ... most recent first
    hist = [np.array([P(t0 - i * dt)]) for i in range(p)]
    S = ScalarState(hist[0].copy(), np.zeros(1), g=lambda tt: dP(tt))
    _, new = hdg_solve_bdf(S, hist, t0 + dt, dt, p)
    assert abs(S.UDG.t[0] - P(t0 + dt)) < 1e-11 * max(1.0, abs(P(t0 + dt)))
    assert S.calls == [(t0 + dt, pytest.approx(bdfcoeff(p)[0] / dt))]
    assert len(new) == p and new[0][0] == S.UDG.t[0]
    for i in range(1, p):
        assert new[i] is hist[i - 1]


def test_bdf_driver_stability():
    """BDF1 on u' = lam u is backward Euler: U1 = U0 / (1 - lam dt)."""
    lam = -np.array([0.0, 1.0, 10.0])
    u0 = np.array([1.0, 2.0, 3.0])
    S = ScalarState(u0, lam)
    hdg_solve_bdf(S, [u0.copy()], 0.1, 0.1, 1)
    assert np.allclose(S.UDG.t, u0 / (1 - lam * 0.1), rtol=1e-14)

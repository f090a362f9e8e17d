"""Pins the oracle's boundary models beyond the far field (the reference's
fbou cases, application/FM/euler/fbou.m:30-91 and application/FM/ns/fbou.m:
47-109): every Jacobian against central finite differences of the value, and
each value by what the physics fixes -- a state that already meets the
boundary condition gives no flux in the constrained rows, a resting gas at
the free-stream temperature is the isothermal wall's reference, and the
prescribed-pressure flux at uh = u is (an - An) acting on the energy
difference alone.  The eps > 0 smoothing of the 2D Euler LLF eigenvalue
(euler/getan.m:97-103) is restated as written, whose derivative lines use the
smoothed value: the test predicts the oracle's Jacobian from the finite
differences and that formula, and checks that both agree where smoothing
leaves the value unchanged."""
import numpy as np
import pytest

from hdg_amd.app import euler_app, ns_app
from tests.test_oracle_physics import _fd, _ns_udg, _states2d

GAM = 1.4


def _pressure(U):
    return (GAM - 1) * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / U[:, 0])


def _normals(n, rng):
    nl = rng.standard_normal((n, 2))
    return nl / np.linalg.norm(nl, axis=1, keepdims=True)


def _case(phys, rng, n=6):
    if phys == "euler":
        app = euler_app(2)
        udg = _states2d(n, rng)
    else:
        app = ns_app(re=100.0)
        udg = _ns_udg(n, rng)
    nl = _normals(n, rng)
    pg = rng.random((n, 2))
    uh = udg[:, :4] * (1 + 0.02 * rng.standard_normal((n, 4)))
    return app, udg, uh, nl, pg


MODELS = [("euler", 2), ("euler", 3), ("euler", 4), ("euler", 5), ("ns", 2), ("ns", 3), ("ns", 4), ("ns", 6)]


@pytest.mark.parametrize("phys,ib", MODELS)
def test_boundary_jacobians_fd(oracle, phys, ib):
    rng = np.random.default_rng(10 + ib + (0 if phys == "euler" else 20))
    app, udg, uh, nl, pg = _case(phys, rng)
    uinf = np.array([1.1, 0.3, -0.2, 3.5])
    fh, fu, fhh = oracle.fbou(app, ib, uinf, nl, pg, udg, uh)
    fd_u = _fd(lambda x: oracle.fbou(app, ib, uinf, nl, pg, x, uh)[0], udg)
    fd_h = _fd(lambda x: oracle.fbou(app, ib, uinf, nl, pg, udg, x)[0], uh)
    assert np.abs(fu - fd_u).max() < 1e-6 * max(1, np.abs(fu).max())
    assert np.abs(fhh - fd_h).max() < 1e-6 * max(1, np.abs(fhh).max())


def test_euler_wall_and_inflow_values(oracle):
    rng = np.random.default_rng(3)
    app, udg, uh, nl, pg = _case("euler", rng)
    # a tangential state (zero normal velocity) with uh = u meets the wall
    un = udg[:, 1] * nl[:, 0] + udg[:, 2] * nl[:, 1]
    ut = udg.copy()
    ut[:, 1] -= un * nl[:, 0]
    ut[:, 2] -= un * nl[:, 1]
    fz, _, _ = oracle.fbou(app, 2, np.zeros(4), nl, pg, ut, ut)
    assert np.abs(fz).max() < 1e-13
    # and the wall reflects only the normal momentum: fh = u - (ru.n) n - uh
    fh, _, _ = oracle.fbou(app, 2, np.zeros(4), nl, pg, udg, uh)
    assert np.allclose(fh[:, 1:3], udg[:, 1:3] - un[:, None] * nl - uh[:, 1:3], atol=1e-14)
    # inflow: density and energy from u, momentum prescribed
    uinf = np.array([9.0, 0.7, -0.4, 9.0])
    fh, _, _ = oracle.fbou(app, 5, uinf, nl, pg, udg, uh)
    assert np.allclose(fh[:, [0, 3]], udg[:, [0, 3]] - uh[:, [0, 3]], atol=1e-14)
    assert np.allclose(fh[:, 1:3], uinf[1:3] - uh[:, 1:3], atol=1e-14)


@pytest.mark.parametrize("ib", [3, 4])
def test_euler_prescribed_pressure_value(oracle, ib):
    """At uh = u the far-field form leaves (an - An)(u - ub); ub differs from u
    only in the energy, by (p(u) - pb)/(gam - 1), and the LLF pair is diagonal
    with an - An = |un| + c - (un - c)."""
    rng = np.random.default_rng(40 + ib)
    app, udg, _, nl, pg = _case("euler", rng)
    uinf = np.array([2.2, 0.0, 0.0, 5.9])
    pb = uinf[0] if ib == 3 else (uinf[3] - 0.5) * (GAM - 1)
    fh, _, _ = oracle.fbou(app, ib, uinf, nl, pg, udg, udg[:, :4])
    r = udg[:, 0]
    un = (udg[:, 1] * nl[:, 0] + udg[:, 2] * nl[:, 1]) / r
    c = np.sqrt(GAM * _pressure(udg) / r)
    expect = (np.abs(un) - un + 2 * c) * (_pressure(udg) - pb) / (GAM - 1)
    assert np.abs(fh[:, :3]).max() < 1e-13
    assert np.allclose(fh[:, 3], expect, rtol=1e-13, atol=1e-13)
    # a state at the prescribed pressure is its own boundary state: no flux
    U = udg[:, :4].copy()
    U[:, 3] += (pb - _pressure(U)) / (GAM - 1)
    fz, _, _ = oracle.fbou(app, ib, uinf, nl, pg, U, U)
    assert np.abs(fz).max() < 1e-12


def test_ns_prescribed_pressure_at_the_pressure(oracle):
    rng = np.random.default_rng(7)
    app, udg, _, nl, pg = _case("ns", rng)
    uinf = np.array([0.0, 0.0, 0.0, 8.0])
    pb = (uinf[3] - 0.5) * (GAM - 1)
    U = udg.copy()
    U[:, 3] += (pb - _pressure(U)) / (GAM - 1)
    fz, _, _ = oracle.fbou(app, 4, uinf, nl, pg, U, U[:, :4])
    assert np.abs(fz).max() < 1e-12


def test_ns_walls_values(oracle):
    rng = np.random.default_rng(8)
    app, udg, uh, nl, pg = _case("ns", rng)
    minf = app.param[4]
    # adiabatic: density from u, no-slip momentum, the interior fhat's energy row
    fh, fu, fhh = oracle.fbou(app, 2, np.zeros(4), nl, pg, udg, uh)
    fa, fau, fah = oracle.fhat(app, nl, pg, udg, uh)
    assert np.allclose(fh[:, 0], udg[:, 0] - uh[:, 0], atol=1e-14)
    assert np.allclose(fh[:, 1:3], -uh[:, 1:3], atol=1e-14)
    assert np.array_equal(fh[:, 3], fa[:, 3])
    assert np.array_equal(fu[:, 3, :], fau[:, 3, :]) and np.array_equal(fhh[:, 3, :], fah[:, 3, :])
    # isothermal: gam (gam-1) M^2 E/r is 1 for a resting gas at p = 1/(gam M^2),
    # r = 1 (the free stream's temperature), so uinf(end) = 1 gives no energy row
    rest = np.tile([1.0, 0.0, 0.0, 1.0 / (GAM * minf ** 2) / (GAM - 1)], (udg.shape[0], 1))
    fh, _, _ = oracle.fbou(app, 3, np.array([0.0, 0.0, 0.0, 1.0]), nl, pg, udg, rest)
    assert np.abs(fh[:, 3]).max() < 1e-13
    # at twice that temperature the row reads 1
    hot = rest.copy()
    hot[:, 3] *= 2.0
    fh, _, _ = oracle.fbou(app, 3, np.array([0.0, 0.0, 0.0, 1.0]), nl, pg, udg, hot)
    assert np.allclose(fh[:, 3], 1.0, atol=1e-13)
    # Dirichlet
    uinf = np.array([1.0, 0.4, 0.1, 6.0])
    fh, fu, fhh = oracle.fbou(app, 6, uinf, nl, pg, udg, uh)
    assert np.allclose(fh, uh - uinf, atol=1e-14) and np.abs(fu).max() == 0.0


def _llf(U, nl):
    r = U[:, 0]
    un = (U[:, 1] * nl[:, 0] + U[:, 2] * nl[:, 1]) / r
    c = np.sqrt(GAM * _pressure(U) / r)
    return np.abs(un) + c, un - c, c


def test_euler_eps_far_field_as_written(oracle):
    rng = np.random.default_rng(11)
    eps = 0.7
    app = euler_app(2, eps=eps)
    n = 6
    udg = _states2d(n, rng)
    nl = _normals(n, rng)
    pg = rng.random((n, 2))
    uh = udg * (1 + 0.02 * rng.standard_normal((n, 4)))
    uinf = app.bcs[0]
    fh, fu, fhh = oracle.fbou(app, 1, uinf, nl, pg, udg, uh)
    # value: an = 0.5 (l^2/(eps c) + eps c), An = un - c, both diagonal
    l, l3, c = _llf(uh, nl)
    an = 0.5 * (l * l / (eps * c) + eps * c)
    expect = an[:, None] * (udg - uinf) + l3[:, None] * (udg + uinf - 2 * uh)
    assert np.allclose(fh, expect, rtol=1e-13, atol=1e-13)
    assert np.allclose(fu, (an + l3)[:, None, None] * np.eye(4)[None], atol=1e-14)
    # d/duh: the FD Jacobian takes the chain rule; the reference's lines differ
    # in an's derivative by (an - l) l'/(eps c) - 0.5 (an^2 - l^2)/(eps c)^2 eps c'
    dl = _fd(lambda x: _llf(x, nl)[0], uh)
    dc = _fd(lambda x: _llf(x, nl)[2], uh)
    ec = (eps * c)[:, None]
    quirk = (an[:, None] - l[:, None]) * dl / ec - 0.5 * (an * an - l * l)[:, None] / ec ** 2 * eps * dc
    fd_h = _fd(lambda x: oracle.fbou(app, 1, uinf, nl, pg, udg, x)[0], uh)
    predicted = fd_h + (udg - uinf)[:, :, None] * quirk[:, None, :]
    assert np.abs(fhh - predicted).max() < 1e-6 * max(1, np.abs(fhh).max())
    assert np.abs(quirk).max() > 1e-3          # the two forms do differ here
    # where the smoothing leaves l unchanged (eps c = l) they coincide
    eps1 = float(l[0] / c[0])
    app1 = euler_app(2, eps=eps1)
    _, _, h1 = oracle.fbou(app1, 1, uinf, nl[:1], pg[:1], udg[:1], uh[:1])
    fd1 = _fd(lambda x: oracle.fbou(app1, 1, uinf, nl[:1], pg[:1], udg[:1], x)[0], uh[:1])
    assert np.abs(h1 - fd1).max() < 1e-6 * max(1, np.abs(h1).max())


def test_unrestated_models_are_refused(oracle):
    rng = np.random.default_rng(2)
    app, udg, uh, nl, pg = _case("ns", rng)
    for ib in (5, 7, 8):
        with pytest.raises(RuntimeError):
            oracle.fbou(app, ib, np.zeros(4), nl, pg, udg, uh)

"""Pins the oracle's pointwise physics: Jacobians against central finite
differences, the Euler flux Jacobian against the closed form the reference
writes out (application/FM/euler/flux.m:35-44), the Roe matrix against the
exact normal flux Jacobian and |A|^2 = A^2, and fhat/fbou identities."""
import numpy as np
import pytest

from hdg_amd.app import euler_app, ns_app


def _states2d(n, rng, nc=4):
    r = 1.0 + 0.2 * rng.random(n)
    u = 0.5 + 0.3 * rng.random(n)
    v = 0.3 * rng.standard_normal(n)
    p = 2.0 + rng.random(n)
    rE = p / 0.4 + 0.5 * r * (u * u + v * v)
    return np.stack([r, r * u, r * v, rE], 1)


def _fd(fun, x, h=1e-6):
    """d fun / d x[:, k] for every k by central differences; fun returns (ng, ...)."""
    base = fun(x)
    out = np.zeros(base.shape + (x.shape[1],))
    for k in range(x.shape[1]):
        xp, xm = x.copy(), x.copy()
        s = h * np.maximum(1.0, np.abs(x[:, k]))
        xp[:, k] += s
        xm[:, k] -= s
        out[..., k] = (fun(xp) - fun(xm)) / (2 * s.reshape((-1,) + (1,) * (base.ndim - 1)))
    return out


def test_euler2d_flux_closed_form(oracle):
    rng = np.random.default_rng(0)
    app = euler_app(2)
    U = _states2d(16, rng)
    pg = rng.random((16, 2))
    f, fu = oracle.flux(app, pg, U)
    gam = 1.4
    gam1 = gam - 1
    r, ru, rv, rE = U.T
    uv, vv, E = ru / r, rv / r, rE / r
    af = 0.5 * (uv * uv + vv * vv)
    z, o = np.zeros_like(r), np.ones_like(r)
    # euler/flux.m:36-43 (note the leading minus on every bracket)
    J = np.zeros((16, 4, 2, 4))
    J[:, :, 0, 0] = -np.stack([z, 0.5 * ((3 - gam) * uv * uv - gam1 * vv * vv), uv * vv, gam * E * uv - 2 * gam1 * uv * af], 1)
    J[:, :, 0, 1] = -np.stack([-o, (gam - 3) * uv, -vv, -gam * E + 0.5 * gam1 * (3 * uv * uv + vv * vv)], 1)
    J[:, :, 0, 2] = -np.stack([z, gam1 * vv, -uv, gam1 * uv * vv], 1)
    J[:, :, 0, 3] = -np.stack([z, -gam1 * o, z, -gam * uv], 1)
    J[:, :, 1, 0] = -np.stack([z, uv * vv, 0.5 * ((3 - gam) * vv * vv - gam1 * uv * uv), gam * E * vv - 2 * gam1 * vv * af], 1)
    J[:, :, 1, 1] = -np.stack([z, -vv, gam1 * uv, gam1 * uv * vv], 1)
    J[:, :, 1, 2] = -np.stack([-o, -uv, (gam - 3) * vv, -gam * E + 0.5 * gam1 * (3 * vv * vv + uv * uv)], 1)
    J[:, :, 1, 3] = -np.stack([z, z, -gam1 * o, -gam * vv], 1)
    assert np.abs(fu - J).max() < 1e-12 * np.abs(J).max()


def test_euler3d_flux_fd(oracle):
    rng = np.random.default_rng(1)
    app = euler_app(3)
    U2 = _states2d(10, rng)
    w = 0.2 * rng.standard_normal(10)
    U = np.stack([U2[:, 0], U2[:, 1], U2[:, 2], U2[:, 0] * w, U2[:, 3] + 0.5 * U2[:, 0] * w * w], 1)
    pg = rng.random((10, 3))
    f, fu = oracle.flux(app, pg, U)
    fd = _fd(lambda x: oracle.flux(app, pg, x)[0], U)
    assert np.abs(fu - fd).max() < 1e-6 * np.abs(fu).max()
    # mass flux is the momentum
    assert np.abs(f[:, 0, :] - U[:, 1:4]).max() < 1e-15


def _ns_udg(n, rng):
    U = _states2d(n, rng)
    q = 0.3 * rng.standard_normal((n, 8))
    return np.hstack([U, q])


@pytest.mark.parametrize("av", [False, True])
def test_ns2d_flux_fd(oracle, av):
    rng = np.random.default_rng(2)
    app = ns_app(re=100.0, av=av)
    udg = _ns_udg(12, rng)
    pg = np.hstack([rng.random((12, 2)), 1e-2 * rng.random((12, 1))])
    f, fu = oracle.flux(app, pg if av else pg[:, :2], udg)
    fd = _fd(lambda x: oracle.flux(app, pg if av else pg[:, :2], x)[0], udg)
    assert np.abs(fu - fd).max() < 1e-6 * np.abs(fu).max()
    if av:
        f0, _ = oracle.flux(ns_app(re=100.0), pg[:, :2], udg)
        q = udg[:, 4:].reshape(12, 4, 2, order="F")
        assert np.abs(f - f0 - pg[:, 2:3, None] * q).max() < 1e-14


def test_ns2d_flux_inviscid_limit(oracle):
    """With q = 0 the NS flux is the Euler flux."""
    rng = np.random.default_rng(3)
    U = _states2d(8, rng)
    udg = np.hstack([U, np.zeros((8, 8))])
    pg = rng.random((8, 2))
    fns, _ = oracle.flux(ns_app(), pg, udg)
    feu, _ = oracle.flux(euler_app(2), pg, U)
    assert np.abs(fns - feu).max() < 1e-13


def _ns_textbook(gam, re, pr, prim, grad):
    """Compressible Navier-Stokes flux in the nondimensional form of
    application/FM/ns/flux2d.m, written from primitive fields: the Euler flux
    minus the viscous flux of a Newtonian fluid (Stokes' hypothesis) with
    Fourier heat conduction in T = p/rho,
      tau = (1/Re) (grad v + grad v^T - 2/3 div v I),
      q_heat = gam / ((gam-1) Re Pr) grad(p/rho).
    prim (n, 4) = rho, u, v, p; grad (n, 4, 2) their x, y derivatives."""
    r, u, v, p = prim.T
    (rx, ry), (ux, uy), (vx, vy), (px, py) = (grad[:, i, :].T for i in range(4))
    rE = p / (gam - 1) + 0.5 * r * (u * u + v * v)
    mu = 1.0 / re
    txx = mu * (2 * ux - (2.0 / 3.0) * (ux + vy))
    tyy = mu * (2 * vy - (2.0 / 3.0) * (ux + vy))
    txy = mu * (uy + vx)
    kappa = gam / ((gam - 1) * re * pr)
    Tx = (px * r - p * rx) / r ** 2
    Ty = (py * r - p * ry) / r ** 2
    f = np.zeros((prim.shape[0], 4, 2))
    f[:, :, 0] = np.stack([r * u, r * u * u + p - txx, r * u * v - txy, u * (rE + p) - (u * txx + v * txy + kappa * Tx)], 1)
    f[:, :, 1] = np.stack([r * v, r * u * v - txy, r * v * v + p - tyy, v * (rE + p) - (u * txy + v * tyy + kappa * Ty)], 1)
    return f


def _ns_udg_from_primitive(gam, prim, grad):
    """udg = (U, q) with U the conservative state and q = -grad U (the HDG
    convention q + grad u = 0, kernel/volint.m:57-60), by the product rule."""
    r, u, v, p = prim.T
    U = np.stack([r, r * u, r * v, p / (gam - 1) + 0.5 * r * (u * u + v * v)], 1)
    q = np.zeros((prim.shape[0], 4, 2))
    for d in range(2):
        rd, ud, vd, pd = (grad[:, i, d] for i in range(4))
        q[:, 0, d] = -rd
        q[:, 1, d] = -(u * rd + r * ud)
        q[:, 2, d] = -(v * rd + r * vd)
        q[:, 3, d] = -(pd / (gam - 1) + 0.5 * (u * u + v * v) * rd + r * (u * ud + v * vd))
    return np.hstack([U, q.reshape(-1, 8, order="F")])


def test_ns2d_flux_closed_form(oracle):
    """The oracle's viscous stress and heat flux against the textbook
    Navier-Stokes flux (application/FM/ns/flux2d.m:107 generates the same
    expressions symbolically).  Re = 10 makes the viscous part a tenth of the
    flux, so a swapped 2/3 <-> 4/3 in tau, a missing gam in the conduction
    coefficient or a sign slip in tau_xy changes f by far more than the bar."""
    rng = np.random.default_rng(17)
    n, gam, re, pr = 24, 1.4, 10.0, 0.72
    prim = np.stack([1.0 + 0.3 * rng.random(n), 0.5 + 0.5 * rng.random(n),
                     0.4 * rng.standard_normal(n), 2.0 + rng.random(n)], 1)
    grad = rng.standard_normal((n, 4, 2))
    udg = _ns_udg_from_primitive(gam, prim, grad)
    f, _ = oracle.flux(ns_app(gam=gam, re=re, pr=pr), rng.random((n, 2)), udg)
    ref = _ns_textbook(gam, re, pr, prim, grad)
    assert np.abs(f - ref).max() < 1e-12 * np.abs(ref).max()
    # the viscous part is large enough for the planted mistakes to show
    inv = _ns_textbook(gam, 1e300, pr, prim, grad)
    assert np.abs(ref - inv).max() > 1e-3 * np.abs(ref).max()


@pytest.mark.parametrize("mistake", ["c23", "gamma", "txy_sign"])
def test_ns2d_flux_closed_form_catches_mistakes(oracle, mistake):
    """The closed-form check fails for each plausible slip in the viscous
    terms: planted in the textbook side (on the CPU), never in the product."""
    rng = np.random.default_rng(18)
    n, gam, re, pr = 24, 1.4, 10.0, 0.72
    prim = np.stack([1.0 + 0.3 * rng.random(n), 0.5 + 0.5 * rng.random(n),
                     0.4 * rng.standard_normal(n), 2.0 + rng.random(n)], 1)
    grad = rng.standard_normal((n, 4, 2))
    udg = _ns_udg_from_primitive(gam, prim, grad)
    f, _ = oracle.flux(ns_app(gam=gam, re=re, pr=pr), rng.random((n, 2)), udg)
    ref = _ns_textbook(gam, re, pr, prim, grad)
    r, u, v, p = prim.T
    (rx, ry), (ux, uy), (vx, vy), (px, py) = (grad[:, i, :].T for i in range(4))
    bad = ref.copy()
    if mistake == "c23":      # tau_xx with 2/3 and 4/3 swapped
        d = (1.0 / re) * ((2.0 / 3.0 * ux - 4.0 / 3.0 * vy) - (4.0 / 3.0 * ux - 2.0 / 3.0 * vy))
        bad[:, 1, 0] -= d
        bad[:, 3, 0] -= u * d
    elif mistake == "gamma":  # conduction coefficient without gam
        Tx = (px * r - p * rx) / r ** 2
        bad[:, 3, 0] += (gam - 1.0) / ((gam - 1) * re * pr) * Tx
    else:                     # tau_xy with the wrong sign
        txy = (1.0 / re) * (uy + vx)
        bad[:, 2, 0] += 2 * txy
        bad[:, 1, 1] += 2 * txy
    assert np.abs(f - bad).max() > 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("which", ["euler2d", "ns2d", "euler3d"])
def test_fhat_jacobians_fd(oracle, which):
    rng = np.random.default_rng(4)
    n = 9
    if which == "euler3d":
        app = euler_app(3)
        U2 = _states2d(n, rng)
        w = 0.1 * rng.standard_normal(n)
        udg = np.stack([U2[:, 0], U2[:, 1], U2[:, 2], U2[:, 0] * w, U2[:, 3] + 0.5 * U2[:, 0] * w * w], 1)
        nl = rng.standard_normal((n, 3))
        pg = rng.random((n, 3))
    elif which == "ns2d":
        app = ns_app(re=100.0)
        udg = _ns_udg(n, rng)
        nl = rng.standard_normal((n, 2))
        pg = rng.random((n, 2))
    else:
        app = euler_app(2)
        udg = _states2d(n, rng)
        nl = rng.standard_normal((n, 2))
        pg = rng.random((n, 2))
    nl /= np.linalg.norm(nl, axis=1, keepdims=True)
    ncu = app.ncu
    uh = udg[:, :ncu] * (1 + 0.01 * rng.standard_normal((n, ncu)))
    fh, fu, fhh = oracle.fhat(app, nl, pg, udg, uh)
    fd_u = _fd(lambda x: oracle.fhat(app, nl, pg, x, uh)[0], udg)
    fd_h = _fd(lambda x: oracle.fhat(app, nl, pg, udg, x)[0], uh)
    assert np.abs(fu - fd_u).max() < 1e-6 * max(1, np.abs(fu).max())
    assert np.abs(fhh - fd_h).max() < 1e-6 * max(1, np.abs(fhh).max())
    # consistency: at u = uh the numerical flux is the physical normal flux
    fh0, _, _ = oracle.fhat(app, nl, pg, np.hstack([uh, udg[:, ncu:]]), uh)
    f, _ = oracle.flux(app, pg, np.hstack([uh, udg[:, ncu:]]))
    assert np.abs(fh0 - np.einsum("gcd,gd->gc", f, nl)).max() < 1e-12 * np.abs(f).max()


@pytest.mark.parametrize("which", ["euler2d", "ns2d", "euler3d"])
def test_fbou_farfield(oracle, which):
    rng = np.random.default_rng(5)
    n = 7
    if which == "euler3d":
        app = euler_app(3)
        U2 = _states2d(n, rng)
        w = 0.1 * rng.standard_normal(n)
        udg = np.stack([U2[:, 0], U2[:, 1], U2[:, 2], U2[:, 0] * w, U2[:, 3] + 0.5 * U2[:, 0] * w * w], 1)
        nd = 3
    elif which == "ns2d":
        app = ns_app(re=100.0)
        udg = _ns_udg(n, rng)
        nd = 2
    else:
        app = euler_app(2)
        udg = _states2d(n, rng)
        nd = 2
    ncu = app.ncu
    nl = rng.standard_normal((n, nd))
    nl /= np.linalg.norm(nl, axis=1, keepdims=True)
    pg = rng.random((n, nd))
    uh = udg[:, :ncu] * (1 + 0.01 * rng.standard_normal((n, ncu)))
    uinf = app.bcs[0]
    fh, fu, fhh = oracle.fbou(app, 1, uinf, nl, pg, udg, uh)
    fd_u = _fd(lambda x: oracle.fbou(app, 1, uinf, nl, pg, x, uh)[0], udg)
    fd_h = _fd(lambda x: oracle.fbou(app, 1, uinf, nl, pg, udg, x)[0], uh)
    assert np.abs(fu - fd_u).max() < 1e-6 * max(1, np.abs(fu).max())
    assert np.abs(fhh - fd_h).max() < 1e-5 * max(1, np.abs(fhh).max())
    # u = uh = uinf gives zero boundary flux
    ui = np.tile(uinf[:ncu], (n, 1))
    fz, _, _ = oracle.fbou(app, 1, uinf, nl, pg, np.hstack([ui, udg[:, ncu:]]), ui)
    assert np.abs(fz).max() < 1e-12


@pytest.mark.parametrize("nd", [2, 3])
def test_roe_matrix_identities(oracle, nd):
    """The far-field flux with u = uh, uinf - uh = e_k exposes the columns of the
    Roe matrices: fh = -(an - An) e_k.  Check an = dF.n/dU (the exact normal
    flux Jacobian) and An^2 = an^2 (|A| = R|L|R^-1) with eps = 0."""
    rng = np.random.default_rng(6)
    if nd == 2:
        # ns far field uses the Roe form (ns/fbou.m:37-46); eps=0 disables the entropy fix
        app = ns_app(eps=0.0, re=100.0)
    else:
        app = euler_app(3)
    ncu = app.ncu
    U = _states2d(1, rng)[0] if nd == 2 else np.array([1.1, 0.6, 0.2, -0.1, 6.0])
    n = rng.standard_normal(nd)
    n /= np.linalg.norm(n)
    nl = n[None, :]
    pg = np.zeros((1, nd))
    udg = np.zeros((1, app.nc))
    udg[0, :ncu] = U
    # fh_udg = an + An; column-probe (an - An) through uinf
    _, fu, _ = oracle.fbou(app, 1, U, nl, pg, udg, U[None, :])
    apA = fu[0, :, :ncu]
    amA = np.zeros((ncu, ncu))
    for k in range(ncu):
        uinf = U.copy()
        uinf[k] += 1.0
        fh, _, _ = oracle.fbou(app, 1, uinf, nl, pg, udg, U[None, :])
        amA[:, k] = -fh[0]
    an = 0.5 * (apA + amA)
    An = 0.5 * (apA - amA)
    _, fu_flux = oracle.flux(app, pg, udg)
    A = np.einsum("cdk,d->ck", fu_flux[0, :, :, :ncu], n)
    assert np.abs(an - A).max() < 1e-12 * np.abs(A).max()
    assert np.abs(An @ An - A @ A).max() < 1e-11 * np.abs(A @ A).max()
    # |A| is positive semi-definite in the sense of its eigenvalues
    assert np.all(np.linalg.eigvals(An).real > -1e-12)

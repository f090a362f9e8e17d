"""Seeded synthetic inputs for the elemental path: meshes, states and traces.

Builds (master, app, mesh, UDG, UH, SH) for the workloads the library is
measured on.  States follow ``initu`` (util/initu.m: nodal values of a field
evaluated at dgnodes) and traces follow ``inituhat`` (util/inituhat.m:17-27:
each trace node gets the average of the adjacent elements' traces), gathered
per element as in kernel/hdg_assemble.m:34.
"""
from dataclasses import dataclass

import numpy as np

from .app import App, euler_app, ns_app
from .master import Master, mkmaster
from .mesh import Mesh, mkmesh_cube, mkmesh_square


@dataclass
class Problem:
    master: Master
    app: App
    mesh: Mesh
    udg: np.ndarray      # (npv, nc, ne)
    uh: np.ndarray       # (ncu, npf*nfe, ne) gathered
    sh: np.ndarray       # (npv, nc, ne)
    name: str = ""
    UH: np.ndarray = None  # (ncu, nsiz) global trace (uh is UH gathered per element)

    @property
    def ne(self):
        return self.mesh.ne


def _rng(seed):
    return np.random.default_rng(np.random.SeedSequence(int(seed)))


def inituhat(master, elcon, udg, ncu, nsiz):
    """Global trace values UH (ncu, nsiz): average of element traces
    (util/inituhat.m:17-27)."""
    perm = master.perm
    npf, nfe = perm.shape
    ne = udg.shape[2]
    fvals = udg[perm.ravel(order="F"), :ncu, :]                 # (npf*nfe, ncu, ne)
    idx = elcon.ravel(order="F")                                 # (npf*nfe*ne)
    # elcon(:, e) lists the element's face nodes in the same (k + npf*j) order
    vals = np.transpose(fvals, (0, 2, 1))                        # (npf*nfe, ne, ncu)
    vals = vals.reshape(npf * nfe * ne, ncu, order="F")
    UH = np.zeros((nsiz, ncu))
    np.add.at(UH, idx, vals)
    cnt = np.bincount(idx, minlength=nsiz).astype(np.float64)
    return (UH / cnt[:, None]).T.copy(order="F")


def gather_uh(UH, elcon):
    """uh = UH(:, elcon) reshaped (ncu, npf*nfe, ne) (kernel/hdg_assemble.m:34)."""
    ncu = UH.shape[0]
    g = UH[:, elcon.ravel(order="F")]
    return np.asfortranarray(g.reshape(ncu, elcon.shape[0], elcon.shape[1], order="F"))


def _bumps(x, centres, sigma, amp):
    g = np.zeros(x.shape[:-1])
    for c in centres:
        g += np.exp(-np.sum((x - c) ** 2, axis=-1) / sigma ** 2)
    return amp * g


def euler_state(dgnodes, nd, gam, minf, seed, nbumps=3, amp=0.01, sigma=0.1, bump_p=True, alpha=0.0):
    """Free stream M at angle alpha (x-y plane), plus Gaussian bumps of relative
    amplitude amp in rho (and p), centres U(0.2, 0.8)^nd, width sigma, from the
    seed."""
    rng = _rng(seed)
    X = np.transpose(dgnodes[:, :nd, :], (0, 2, 1))              # (npv, ne, nd)
    centres = rng.uniform(0.2, 0.8, size=(nbumps, nd))
    g = _bumps(X, centres, sigma, amp)
    pinf = 1.0 / (gam * minf ** 2)
    r = 1.0 + g
    p = pinf * (1.0 + g) if bump_p else np.full_like(g, pinf)
    vel = [np.full_like(g, np.cos(alpha)), np.full_like(g, np.sin(alpha))] + [np.zeros_like(g)] * (nd - 2)
    rE = p / (gam - 1.0) + 0.5 * r * sum(v * v for v in vel)
    comps = [r] + [r * v for v in vel] + [rE]
    U = np.stack(comps, axis=1)                                  # (npv, ncu, ne)
    return np.asfortranarray(U)


def ns_state(dgnodes, gam, minf, seed, amp=0.02):
    """Free stream M with a random-phase sinusoidal velocity perturbation of
    relative amplitude amp; q = -grad(u) of the conservative field (the HDG
    convention of kernel/volint.m:57-60, where Rq = 0 for q = -grad u)."""
    rng = _rng(seed)
    ph = rng.uniform(0.0, 1.0, size=4)
    x = dgnodes[:, 0, :]
    y = dgnodes[:, 1, :]
    tp = 2.0 * np.pi
    sx, cx = np.sin(tp * (x + ph[0])), np.cos(tp * (x + ph[0]))
    sy, cy = np.sin(tp * (y + ph[1])), np.cos(tp * (y + ph[1]))
    s2x, c2x = np.sin(tp * (x + ph[2])), np.cos(tp * (x + ph[2]))
    s2y, c2y = np.sin(tp * (y + ph[3])), np.cos(tp * (y + ph[3]))
    u = 1.0 + amp * sx * sy
    v = amp * s2x * c2y
    ux, uy = amp * tp * cx * sy, amp * tp * sx * cy
    vx, vy = amp * tp * c2x * c2y, -amp * tp * s2x * s2y
    pinf = 1.0 / (gam * minf ** 2)
    r = np.ones_like(u)
    rE = pinf / (gam - 1.0) + 0.5 * (u * u + v * v)
    rEx = u * ux + v * vx
    rEy = u * uy + v * vy
    zero = np.zeros_like(u)
    comps = [r, u, v, rE, -zero, -ux, -vx, -rEx, -zero, -uy, -vy, -rEy]
    return np.asfortranarray(np.stack(comps, axis=1))


def _finish(master, app, mesh, udg, name, tdep_sh=None):
    uhg = inituhat(master, mesh.elcon, udg, app.ncu, mesh.nsiz)
    uh = gather_uh(uhg, mesh.elcon)
    sh = np.zeros_like(udg) if tdep_sh is None else tdep_sh
    return Problem(master=master, app=app, mesh=mesh, udg=np.asfortranarray(udg), uh=uh,
                   sh=np.asfortranarray(sh), name=name, UH=uhg)


def euler2d_square(n, porder, seed=0, minf=0.5, jitter_amp=0.1, alpha=0.0):
    """Euler 2D on a jittered 2*n*n triangle square (the 2x32x32, p=2 workload
    at n=32)."""
    m = mkmaster(2, porder)
    app = euler_app(2, minf=minf, alpha=alpha)
    mesh = mkmesh_square(n, n, m, jitter_amp, seed)
    udg = euler_state(mesh.dgnodes, 2, 1.4, minf, seed + 1, alpha=alpha)
    return _finish(m, app, mesh, udg, f"euler2d_p{porder}_n{n}")


def ns2d_square(n, porder=3, seed=0, minf=0.3, av_frac=0.05, av_max=1e-3, jitter_amp=0.1):
    """NS 2D (u, q) on a jittered 2*n*n triangle square with a laplacian AV
    field of av_max on a seeded av_frac of the elements (2x128x128, p=3 at
    n=128)."""
    m = mkmaster(2, porder)
    app = ns_app(minf=minf, av=True)
    mesh = mkmesh_square(n, n, m, jitter_amp, seed, ncd_extra=1)
    rng = _rng(seed + 7)
    hot = rng.random(mesh.ne) < av_frac
    mesh.dgnodes[:, 2, :] = np.where(hot[None, :], av_max, 0.0)
    udg = ns_state(mesh.dgnodes, 1.4, minf, seed + 1)
    return _finish(m, app, mesh, udg, f"ns2d_p{porder}_n{n}")


def euler3d_cube(n, porder=2, seed=0, minf=0.4, jitter_amp=0.1, alpha=0.0):
    """Euler 3D on a jittered 6*n^3 tetrahedron cube with 1% density bumps
    (6x40^3, p=2 at n=40)."""
    m = mkmaster(3, porder)
    app = euler_app(3, minf=minf, alpha=alpha)
    mesh = mkmesh_cube(n, m, jitter_amp, seed)
    udg = euler_state(mesh.dgnodes, 3, 1.4, minf, seed + 1, bump_p=False, alpha=alpha)
    return _finish(m, app, mesh, udg, f"euler3d_p{porder}_n{n}")


def with_bdf1(P, dt, seed=0, amp=0.01):
    """Turns P into one BDF1 step (kernel/hdg_solve_bdf.m:29-43, torder 1):
    tdep = 1, fc_u = 1/dt, SH = u_n/dt with the previous state u_n a seeded
    perturbation of P.udg (q components of SH zero)."""
    rng = _rng(seed)
    ncu = P.app.ncu
    un = P.udg[:, :ncu, :] * (1.0 + amp * rng.standard_normal(P.udg[:, :ncu, :].shape))
    sh = np.zeros_like(P.udg)
    sh[:, :ncu, :] = un / dt
    P.app.tdep = 1
    P.app.fc_u = 1.0 / dt
    P.sh = np.asfortranarray(sh)
    P.name += "_bdf1"
    return P


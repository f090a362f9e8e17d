"""Pins the getq restatement (oracle/getq_oracle.py, kernel/getq.m) without the
reference:

* closed form: on straight-sided elements, with u a polynomial of degree <= p
  in x and uh its trace, the discrete q equals -grad u / fc_q exactly, and the
  nodal gradient qq equals -grad u at the nodes;
* it zeroes the local residual of the C++ restatement of volint/faceint:
  Rq(u, q, uh) = M (SH_q/fc_q - q) + C u - E uh = 0 (volint.m:57-60,
  faceint.m:48-51) with fc_q != 1 and a nonzero SH;
* a wrong sign of either getq.m convention (Xx, or E's -jacf) or a dropped
  face fails both.
"""
import dataclasses

import numpy as np
import pytest

from hdg_amd.master import mkmaster, mkshape
from hdg_amd.mesh import mkmesh_cube, mkmesh_square
from hdg_amd.problems import ns2d_square
from oracle import getq_oracle as G


def poly_field(dgnodes, nd, p, ncu, seed):
    """u_c(x) = a random polynomial of total degree p per component, and its
    gradient, at the element nodes."""
    rng = np.random.default_rng(seed)
    x = dgnodes[:, :nd, :]
    exps = [e for e in np.ndindex(*(p + 1,) * nd) if sum(e) <= p]
    u = np.zeros((x.shape[0], ncu, x.shape[2]))
    gu = np.zeros((x.shape[0], ncu, nd, x.shape[2]))
    for c in range(ncu):
        coef = rng.standard_normal(len(exps))
        for k, e in zip(coef, exps):
            mono = np.prod([x[:, d, :] ** e[d] for d in range(nd)], axis=0)
            u[:, c, :] += k * mono
            for d in range(nd):
                if e[d] > 0:
                    ed = list(e)
                    ed[d] -= 1
                    gu[:, c, d, :] += k * e[d] * np.prod([x[:, i, :] ** ed[i] for i in range(nd)], axis=0)
    return u, gu


def own_trace(u, perm):
    """UHE(c, a + npf*j, e) = u(perm(a, j), c, e): the element's trace of u."""
    npf, nfe = perm.shape
    ne, ncu = u.shape[2], u.shape[1]
    uhe = np.zeros((ncu, npf * nfe, ne))
    for j in range(nfe):
        uhe[:, j * npf:(j + 1) * npf, :] = np.transpose(u[perm[:, j], :, :], (1, 0, 2))
    return np.asfortranarray(uhe)


GEOMS = [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2)]


def _mesh(nd, p, seed=0):
    m = mkmaster(nd, p)
    mesh = mkmesh_square(3, 3, m, jitter_amp=0.15, seed=seed) if nd == 2 else mkmesh_cube(2, m, jitter_amp=0.1, seed=seed)
    return m, mesh


@pytest.mark.parametrize("nd,p", GEOMS)
def test_getq_closed_form(nd, p):
    m, mesh = _mesh(nd, p, seed=p)
    ncu = nd + 2
    u, gu = poly_field(mesh.dgnodes, nd, p, ncu, seed=10 * nd + p)
    udg = np.asfortranarray(np.concatenate([u, np.zeros((u.shape[0], ncu * nd, u.shape[2]))], axis=1))
    uhe = own_trace(u, np.asarray(m.perm))
    shapnv = mkshape(p, m.plocvl, m.plocvl)
    fc_q = 1.7
    q, qq = G.getq(m, mesh.dgnodes, udg, uhe, fc_q=fc_q, shapnv=shapnv)
    want = -gu.reshape(u.shape[0], ncu * nd, u.shape[2], order="F")
    scale = np.abs(want).max()
    assert np.abs(q * fc_q - want).max() <= 1e-11 * scale
    assert np.abs(qq - want).max() <= 1e-11 * scale
    # SH shifts q by SH_q/fc_q
    sh = np.random.default_rng(1).standard_normal(udg.shape)
    q2 = G.getq(m, mesh.dgnodes, udg, uhe, SH=sh, fc_q=fc_q)
    assert np.abs(q2 - q - sh[:, ncu:, :] / fc_q).max() <= 1e-12 * scale


@pytest.mark.parametrize("p", [2, 3])
def test_getq_zeroes_rq(oracle, p):
    P = ns2d_square(2, p, seed=p)
    app = dataclasses.replace(P.app, fc_q=2.5)
    ncu, nd = app.ncu, 2
    rng = np.random.default_rng(p)
    sh = np.zeros_like(P.udg)
    sh[:, ncu:, :] = 0.1 * rng.standard_normal(sh[:, ncu:, :].shape)
    q = G.getq(P.master, P.mesh.dgnodes, P.udg, P.uh, SH=sh, fc_q=app.fc_q)
    udg = P.udg.copy(order="F")
    udg[:, ncu:, :] = q
    r = oracle.elemental_matrices(P.master, app, P.mesh.dgnodes, P.mesh.bf, udg, P.uh, sh)["Rq"]
    r0 = oracle.elemental_matrices(P.master, app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, sh)["Rq"]
    assert np.abs(r0).max() > 1e-3
    assert np.abs(r).max() <= 1e-12 * np.abs(r0).max() + 1e-14

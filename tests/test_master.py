"""Master element: pinned to the reference's golden struct (master/testpython.mat,
p=2 triangle, pgauss=4) and to invariants the mathematics fixes for every
order the kernels are built for."""
import os

import numpy as np
import pytest

from hdg_amd.master import mkmaster, masternodes, mkshape
from hdg_amd.quadrature import tri_rule, tet_rule, line_rule

GOLD = os.path.join(os.path.dirname(__file__), "golden", "master_p2_tri.npz")


def test_golden_p2_triangle():
    g = np.load(GOLD)
    m = mkmaster(2, 2, 4)
    assert int(g["porder"][0]) == 2 and int(g["pgauss"][0]) == 4
    for k in ["plocvl", "plocfc", "gpvl", "gwvl", "gpfc", "gwfc", "shapvl", "shapfc", "shapvt", "shapvg",
              "shapvgdotshapvl", "shapft", "shapfg", "shapfgdotshapfc"]:
        a = getattr(m, k)
        b = g[k].reshape(a.shape)
        assert np.abs(a - b).max() < 1e-13, k
    assert np.array_equal(m.perm + 1, g["perm"].astype(np.int64))


@pytest.mark.parametrize("nd,p", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)])
def test_shape_invariants(nd, p):
    m = mkmaster(nd, p)
    # nodal: phi_i(x_j) = delta_ij
    S = mkshape(p, m.plocvl, m.plocvl)
    assert np.abs(S[:, :, 0] - np.eye(m.npv)).max() < 1e-12
    # partition of unity and zero-sum derivatives at the Gauss points
    assert np.abs(m.shapvl[:, :, 0].sum(0) - 1).max() < 1e-12
    assert np.abs(m.shapvl[:, :, 1:].sum(0)).max() < 1e-11
    # derivative of x_d interpolant is e_d: sum_i x_i(d) dphi_i/dxi_e = delta_de
    for d in range(nd):
        for e in range(nd):
            v = m.plocvl[:, d] @ m.shapvl[:, :, 1 + e]
            assert np.abs(v - (d == e)).max() < 1e-11
    # the mass matrix is SPD
    M = m.shapvl[:, :, 0] @ np.diag(m.gwvl) @ m.shapvl[:, :, 0].T
    assert np.all(np.linalg.eigvalsh(M) > 0)
    # the face basis reproduces the volume basis traces: phi_perm(a,j) restricted to face j
    assert m.perm.shape == ((p + 1) if nd == 2 else (p + 1) * (p + 2) // 2, nd + 1)


@pytest.mark.parametrize("pg", [2, 4, 6, 8])
def test_triangle_rule_exactness(pg):
    x, w = tri_rule(pg)
    # int_T x^a y^b = a! b! / (a+b+2)!
    from math import factorial
    for a in range(pg + 1):
        for b in range(pg + 1 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert abs((w * x[:, 0] ** a * x[:, 1] ** b).sum() - exact) < 5e-14


@pytest.mark.parametrize("pg", [2, 4, 6])
def test_tet_rule_exactness(pg):
    x, w = tet_rule(pg)
    from math import factorial
    for a in range(pg + 1):
        for b in range(pg + 1 - a):
            for c in range(pg + 1 - a - b):
                exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                # the tetrahedron tables carry 14 digits
                assert abs((w * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c).sum() - exact) < 1e-13


@pytest.mark.parametrize("pg", [2, 4, 6, 8])
def test_line_rule(pg):
    x, w = line_rule(pg)
    assert len(w) == pg // 2 + 1
    for a in range(pg + 1):
        assert abs((w * x[:, 0] ** a).sum() - 1.0 / (a + 1)) < 1e-15


@pytest.mark.parametrize("nd,p", [(2, 1), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)])
def test_face_perm_outward(nd, p):
    """Face nodes of the master element lie on face j (opposite vertex j) and
    their parametrisation gives an outward normal (faceint's normal formula)."""
    plocvl, plocfc, perm = masternodes(p, nd)
    verts = np.vstack([np.zeros(nd), np.eye(nd)])
    cen = verts.mean(0)
    m = mkmaster(nd, p)
    for j in range(nd + 1):
        X = plocvl[perm[:, j] - 1]
        others = np.delete(verts, j, axis=0)
        # all face nodes are in the affine hull of the face's vertices
        A = np.hstack([others, np.ones((nd, 1))]).T
        lam = np.linalg.lstsq(A, np.hstack([X, np.ones((X.shape[0], 1))]).T, rcond=None)[0]
        assert np.abs(A @ lam - np.hstack([X, np.ones((X.shape[0], 1))]).T).max() < 1e-12
        dpg = np.einsum("ga k,ad->gdk".replace(" ", ""), m.shapmf[:, :, 1:], X) if nd == 3 else None
        if nd == 2:
            t = m.shapmf[:, :, 1] @ X
            n = np.stack([t[:, 1], -t[:, 0]], 1)
        else:
            d1 = m.shapmf[:, :, 1] @ X
            d2 = m.shapmf[:, :, 2] @ X
            n = np.cross(d1, d2)
        mid = X.mean(0)
        assert np.all(n @ (mid - cen) > 0)

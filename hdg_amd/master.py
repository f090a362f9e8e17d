"""Master (reference) element: nodes, face permutations, quadrature and nodal
shape functions, plus the precomputed basis products the elemental kernels use.

Mirrors ``master = mkmaster(mesh, pgauss)`` (master/mkmaster.m:1-96) followed by
the master-side half of ``preprocess`` (util/preprocess.m:20-55).  Arrays keep
the reference's MATLAB shapes and are stored column-major (Fortran order), so
``a.ravel(order='F')`` is the memory layout the C ABI expects.

Only simplices (elemtype 0) with uniformly spaced nodes (nodetype 0) are built:
triangles p = 1..4 and tetrahedra p = 1..3.
"""
from dataclasses import dataclass, field

import numpy as np

from .quadrature import line_rule, tri_rule, tet_rule


# ---------------------------------------------------------------------------
# node sets (mesh/mkmesh/masternodes.m)
# ---------------------------------------------------------------------------

def _tri_nodes(p):
    """Uniform nodes of the unit triangle in the reference order: x runs fastest,
    rows of constant y from y=0 upward (masternodes.m trinodes2d, ndgrid order)."""
    pts = [(i / p, j / p) for j in range(p + 1) for i in range(p + 1 - j)]
    return np.array(pts, dtype=np.float64)


def _tri_permedge(p):
    """Face (edge) -> volume node table, 1-based, faces opposite vertices 1,2,3
    (masternodes.m trinodes2d: permedge recurrence)."""
    npv = (p + 1) * (p + 2) // 2
    perm = np.zeros((p + 1, 3), dtype=np.int64)
    perm[0] = [p + 1, npv, 1]
    for ii in range(2, p + 2):
        perm[ii - 1, 0] = perm[ii - 2, 0] + p + 2 - ii
        perm[ii - 1, 1] = perm[ii - 2, 1] - ii
        perm[ii - 1, 2] = ii
    return perm


def _faceindex(p):
    """Re-ordering of a face's nodes used for two tetrahedron faces
    (masternodes.m faceindex)."""
    n = p + 1
    a = np.zeros((n, n), dtype=np.int64)
    k = 1
    for j in range(n):
        for i in range(n - j):
            a[i, j] = k
            k += 1
    b = a.T
    # MATLAB b(b>0) walks b column-major == a row-major
    return np.array([v for v in b.ravel(order="F") if v > 0], dtype=np.int64)


def _tet_nodes(p):
    """Uniform tetrahedron nodes: layers of constant z, each a scaled triangle
    node set (masternodes.m tetnodes3d)."""
    xi = np.linspace(0.0, 1.0, p + 1)
    layers = [np.hstack([_tri_nodes(p), np.zeros((_tri_nodes(p).shape[0], 1))])]
    for ii in range(p - 1, 0, -1):
        t = _tri_nodes(ii)
        layers.append(np.hstack([xi[ii] * t, np.full((t.shape[0], 1), xi[p - ii])]))
    layers.append(np.array([[0.0, 0.0, 1.0]]))
    return np.vstack(layers)


def _tet_permface(p, ploc):
    """Face -> volume node table (1-based) for faces [2,3,4],[1,4,3],[1,2,4],[1,3,2]
    (masternodes.m tetnodes3d: permface)."""
    q = ploc.copy()
    q[np.abs(q) < 1e-7] = 0.0
    fi = _faceindex(p) - 1
    cols = []
    cols.append(np.nonzero(np.abs(q[:, 0] + q[:, 1] + q[:, 2] - 1) < 1e-6)[0])
    cols.append(np.nonzero(np.abs(q[:, 0]) < 1e-6)[0][fi])
    cols.append(np.nonzero(np.abs(q[:, 1]) < 1e-6)[0])
    cols.append(np.nonzero(np.abs(q[:, 2]) < 1e-6)[0][fi])
    return np.stack(cols, axis=1).astype(np.int64) + 1


def masternodes(porder, dim):
    """Returns (plocvl, plocfc, perm) with perm 1-based, shape (npf, nfe)."""
    if porder < 1:
        raise ValueError("porder >= 1 required")
    if dim == 2:
        plocvl = _tri_nodes(porder)
        plocfc = plocvl[: porder + 1, :1].copy()
        perm = _tri_permedge(porder)
    elif dim == 3:
        plocvl = _tet_nodes(porder)
        plocfc = plocvl[: (porder + 1) * (porder + 2) // 2, :2].copy()
        perm = _tet_permface(porder, plocvl)
    else:
        raise ValueError("dim must be 2 or 3")
    return plocvl, plocfc, perm


# ---------------------------------------------------------------------------
# nodal shape functions (master/mkshape.m: nodal basis = modal basis / Vandermonde)
# ---------------------------------------------------------------------------

def _legendre(n, x):
    """Legendre P_0..P_n and derivatives at x in [-1,1]; arrays (n+1, len(x))."""
    P = np.zeros((n + 1, x.size))
    dP = np.zeros((n + 1, x.size))
    P[0] = 1.0
    if n >= 1:
        P[1] = x
        dP[1] = 1.0
    for k in range(1, n):
        P[k + 1] = ((2 * k + 1) * x * P[k] - k * P[k - 1]) / (k + 1)
        dP[k + 1] = dP[k - 1] + (2 * k + 1) * P[k]
    return P, dP


def _modal(porder, pts):
    """Complete degree-p polynomial space on the simplex spanned by products of
    Legendre polynomials in 2x-1 (total degree <= p).  Returns list of arrays
    [f, f_x1, ..., f_xd], each (npts, npoly).  Any basis of the same space gives
    the same nodal functions; this one is orthogonal per coordinate and well
    conditioned for p <= 4."""
    pts = np.atleast_2d(pts)
    npt, dim = pts.shape
    L = [_legendre(porder, 2.0 * pts[:, d] - 1.0) for d in range(dim)]
    idx = []
    if dim == 1:
        idx = [(i,) for i in range(porder + 1)]
    elif dim == 2:
        idx = [(i, j) for s in range(porder + 1) for j in range(s + 1) for i in [s - j]]
    else:
        idx = [(i, j, k) for s in range(porder + 1) for k in range(s + 1)
               for j in range(s - k + 1) for i in [s - j - k]]
    f = np.ones((npt, len(idx)))
    df = [np.ones((npt, len(idx))) for _ in range(dim)]
    for c, ijk in enumerate(idx):
        for d in range(dim):
            P, dP = L[d]
            f[:, c] *= P[ijk[d]]
            for e in range(dim):
                df[e][:, c] *= (2.0 * dP[ijk[d]]) if e == d else P[ijk[d]]
    return [f] + df


def mkshape(porder, plocal, pts):
    """Nodal shape functions and derivatives, shape (np, npoints, dim+1)
    (master/mkshape.m:21-53)."""
    A = _modal(porder, plocal)[0]
    vals = _modal(porder, pts)
    nfs = np.stack([np.linalg.solve(A.T, v.T) for v in vals], axis=2)
    return np.asfortranarray(nfs)


# ---------------------------------------------------------------------------
# master element
# ---------------------------------------------------------------------------

@dataclass
class Master:
    nd: int
    porder: int
    pgauss: int
    plocvl: np.ndarray
    plocfc: np.ndarray
    perm: np.ndarray          # (npf, nfe), 0-based
    permgeom: np.ndarray      # (npf, nfe), 0-based (isoparametric: == perm)
    gpvl: np.ndarray
    gwvl: np.ndarray
    gpfc: np.ndarray
    gwfc: np.ndarray
    shapvl: np.ndarray        # (npv, ngv, nd+1)
    shapfc: np.ndarray        # (npf, ngf, nd)
    shapmv: np.ndarray        # (ngv, npv, nd+1) geometry basis at volume points
    shapmf: np.ndarray        # (ngf, npf, nd)   geometry basis at face points
    shapvt: np.ndarray = field(default=None)
    shapvg: np.ndarray = field(default=None)
    shapvgdotshapvl: np.ndarray = field(default=None)
    shapft: np.ndarray = field(default=None)
    shapfg: np.ndarray = field(default=None)
    shapfgdotshapfc: np.ndarray = field(default=None)

    @property
    def npv(self):
        return self.plocvl.shape[0]

    @property
    def npf(self):
        return self.perm.shape[0]

    @property
    def nfe(self):
        return self.perm.shape[1]

    @property
    def ngv(self):
        return self.gpvl.shape[0]

    @property
    def ngf(self):
        return self.gpfc.shape[0]

    @property
    def ndf(self):
        return self.npf * self.nfe


def mkmaster(nd, porder, pgauss=None):
    """``mkmaster`` + ``preprocess`` master products (master/mkmaster.m:10-96,
    util/preprocess.m:20-55), isoparametric (morder = porder)."""
    if pgauss is None:
        pgauss = max(2 * porder, 1)
    plocvl, plocfc, perm1 = masternodes(porder, nd)
    if nd == 2:
        gpvl, gwvl = tri_rule(pgauss)
        gpfc, gwfc = line_rule(pgauss)
    else:
        gpvl, gwvl = tet_rule(pgauss)
        gpfc, gwfc = tri_rule(pgauss)
    shapvl = mkshape(porder, plocvl, gpvl)
    shapfc = mkshape(porder, plocfc, gpfc)
    m = Master(nd=nd, porder=porder, pgauss=pgauss, plocvl=plocvl, plocfc=plocfc,
               perm=np.asfortranarray(perm1 - 1), permgeom=np.asfortranarray(perm1 - 1),
               gpvl=gpvl, gwvl=gwvl, gpfc=gpfc, gwfc=gwfc, shapvl=shapvl, shapfc=shapfc,
               shapmv=np.asfortranarray(np.transpose(shapvl, (1, 0, 2))),
               shapmf=np.asfortranarray(np.transpose(shapfc, (1, 0, 2))))
    _preprocess_products(m)
    return m


def _preprocess_products(m):
    """shapvt, shapvg = shapvl.*w, shapvgdotshapvl((ii-1)*npv+jj,g,d) =
    shapvg(jj,g,d)*shapvl(ii,g,1); same for faces (util/preprocess.m:23-55)."""
    npv, ngv, nd1 = m.shapvl.shape
    npf, ngf, ndf_ = m.shapfc.shape
    m.shapvt = np.asfortranarray(np.transpose(m.shapvl, (1, 0, 2)))
    m.shapvg = np.asfortranarray(m.shapvl * m.gwvl[None, :, None])
    # index (jj, ii) in column-major npv x npv == linear (ii-1)*npv + jj
    svv = m.shapvg[:, None, :, :] * m.shapvl[None, :, :, 0:1]        # (jj, ii, g, d)
    m.shapvgdotshapvl = np.asfortranarray(svv.reshape(npv * npv, ngv, nd1, order="F"))
    m.shapft = np.asfortranarray(np.transpose(m.shapfc, (1, 0, 2)))
    m.shapfg = np.asfortranarray(m.shapfc * m.gwfc[None, :, None])
    sff = m.shapfg[:, None, :, :] * m.shapfc[None, :, :, 0:1]
    m.shapfgdotshapfc = np.asfortranarray(sff.reshape(npf * npf, ngf, ndf_, order="F"))

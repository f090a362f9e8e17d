"""Structured simplex meshes and the HDG connectivity the elemental path needs.

Mirrors the reference's mesh pipeline for the unit square and cube:
``squaremesh``/``cubemesh`` (mesh/squaremesh.m, mesh/cubemesh.m), face numbering
``mkt2f`` (mesh/mkmesh/mkt2f.m:1-90), boundary tagging ``setbndnbrs``
(mesh/mkmesh/setbndnbrs.m), DG nodes ``createnodes`` (mesh/mkmesh/createnodes.m)
and the trace-node numbering ``hdgelcon`` (mesh/mkmesh/elconnectivities.m:147-212).
Indices are 0-based; array shapes follow the reference (column-major).

Seeded vertex jitter (interior vertices moved by U(-a*h, a*h) per coordinate)
keeps elements straight-sided and positively oriented for a < 0.25.
"""
from dataclasses import dataclass

import numpy as np

# local faces: opposite vertex j for triangles (mkt2f.m face table)
TRI_FACES = np.array([[1, 2], [2, 0], [0, 1]])
TET_FACES = np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]])


@dataclass
class Mesh:
    nd: int
    porder: int
    p: np.ndarray        # (np, nd) vertices
    t: np.ndarray        # (ne, nd+1) 0-based
    f: np.ndarray        # (nf, nvf+2): face vertices, element 1, element 2 (or -bnd id)
    t2f: np.ndarray      # (ne, nfe) 0-based face ids
    bf: np.ndarray       # (nfe, ne): >0 interior (neighbour index+1), <0 boundary -id
    dgnodes: np.ndarray  # (npv, ncd, ne)
    elcon: np.ndarray    # (npf*nfe, ne) 0-based trace node ids
    nsiz: int            # number of trace nodes

    @property
    def ne(self):
        return self.t.shape[0]

    @property
    def nf(self):
        return self.f.shape[0]


def _rng(seed):
    return np.random.default_rng(np.random.SeedSequence(int(seed)))


def squaremesh(m, n, parity=0):
    """Unit-square triangulation with m x n vertices, SW-NE diagonals for
    parity 0, triangles in Cartesian order (mesh/squaremesh.m)."""
    x, y = np.meshgrid(np.arange(m) / (m - 1), np.arange(n) / (n - 1), indexing="ij")
    p = np.stack([x.ravel(order="F"), y.ravel(order="F")], axis=1)
    tris = []
    for j in range(n - 1):
        for i in range(m - 1):
            v = i + j * m
            if parity == 0:
                a = [v, v + 1, v + m + 1]
                b = [v, v + m + 1, v + m]
                tris += [b, a]          # mesh/squaremesh.m reorders as [ix2; ix1]
            else:
                a = [v, v + 1, v + m]
                b = [v + 1, v + m + 1, v + m]
                tris += [a, b]
    return p, np.array(tris, dtype=np.int64)


def cubemesh(m, n, o):
    """Unit-cube tetrahedralisation, 6 tets per cell (mesh/cubemesh.m), with
    every tetrahedron oriented to positive volume."""
    x, y, z = np.meshgrid(np.arange(m) / (m - 1), np.arange(n) / (n - 1),
                          np.arange(o) / (o - 1), indexing="ij")
    p = np.stack([x.ravel(order="F"), y.ravel(order="F"), z.ravel(order="F")], axis=1)
    t0 = np.array([[5, 6, 7, 3], [5, 6, 2, 3], [5, 1, 2, 3], [6, 8, 7, 4], [6, 7, 4, 3],
                   [6, 2, 4, 3]]) - 1
    cmap = np.array([0, 1, m, m + 1, m * n, m * n + 1, m * n + m, m * n + m + 1])
    base = t0.shape[0]
    tets = np.empty(((m - 1) * (n - 1) * (o - 1) * base, 4), dtype=np.int64)
    c = 0
    for k in range(o - 1):
        for j in range(n - 1):
            for i in range(m - 1):
                v = i + j * m + k * m * n
                tets[c:c + base] = cmap[t0] + v
                c += base
    return p, _orient(p, tets)


def _orient(p, t):
    nd = p.shape[1]
    if nd == 3:
        a, b, c, d = (p[t[:, i]] for i in range(4))
        vol = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
        neg = vol < 0
        t = t.copy()
        t[neg, 0], t[neg, 1] = t[neg, 1].copy(), t[neg, 0].copy()
    return t


def jitter(p, h, amp, seed):
    """Move interior vertices by U(-amp*h, amp*h) in each coordinate."""
    rng = _rng(seed)
    interior = np.all((p > 1e-12) & (p < 1 - 1e-12), axis=1)
    q = p.copy()
    q[interior] += rng.uniform(-amp * h, amp * h, size=(interior.sum(), p.shape[1]))
    return q


def mkt2f(t, nd):
    """Face list and element->face map (mesh/mkmesh/mkt2f.m): faces numbered in
    order of first visit over elements and local faces; f rows hold the face's
    vertices as seen from the first element, that element, and the neighbour
    (or -1 for a boundary face, replaced by -bnd id later)."""
    faces = TRI_FACES if nd == 2 else TET_FACES
    ne, nfe = t.shape[0], faces.shape[0]
    fv = t[:, faces].reshape(ne * nfe, -1)          # flat index = e*nfe + j
    key = np.sort(fv, axis=1)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.ravel()
    # face ids in order of first visit over (element, local face)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    t2f = rank[inv]
    owner = first[order]
    nf = order.size
    f = np.empty((nf, fv.shape[1] + 2), dtype=np.int64)
    f[:, :-2] = fv[owner]
    f[:, -2] = owner // nfe
    f[:, -1] = -1
    second = np.ones(ne * nfe, dtype=bool)
    second[owner] = False
    idx2 = np.nonzero(second)[0]
    f[t2f[idx2], -1] = idx2 // nfe
    return f, t2f.reshape(ne, nfe)


def setbndnbrs(p, f, nd):
    """Tags boundary faces with -id: square 1:y=0 2:x=1 3:y=1 4:x=0
    (mesh/mkmesh_square.m bndexpr); cube 1:x=0 2:x=1 3:y=0 4:y=1 5:z=0 6:z=1
    (mesh/mkmesh_cube.m bndexpr)."""
    f = f.copy()
    nvf = nd
    for i in np.nonzero(f[:, -1] < 0)[0]:
        q = p[f[i, :nvf]]
        if nd == 2:
            tests = [np.all(q[:, 1] < 1e-6), np.all(q[:, 0] > 1 - 1e-6),
                     np.all(q[:, 1] > 1 - 1e-6), np.all(q[:, 0] < 1e-6)]
        else:
            tests = [np.all(q[:, 0] < 1e-6), np.all(q[:, 0] > 1 - 1e-6),
                     np.all(q[:, 1] < 1e-6), np.all(q[:, 1] > 1 - 1e-6),
                     np.all(q[:, 2] < 1e-6), np.all(q[:, 2] > 1 - 1e-6)]
        hit = [k for k, ok in enumerate(tests) if ok]
        if not hit:
            raise ValueError("strange boundary face")
        f[i, -1] = -(hit[0] + 1)
    return f


def createnodes(p, t, plocvl):
    """DG nodes by the affine map of each simplex (mesh/mkmesh/createnodes.m)."""
    nd = p.shape[1]
    lam = np.hstack([1.0 - plocvl.sum(axis=1, keepdims=True), plocvl])   # (npv, nd+1)
    X = np.einsum("av,evd->ade", lam, p[t])                              # (npv, nd, ne)
    return np.asfortranarray(X)


def hdgelcon(dgnodes, perm, f, t2f):
    """Trace-node numbering: face i owns nodes i*npf..i*npf+npf-1 in the order of
    its first element; each element's local face nodes are matched to them by
    coordinates (mesh/mkmesh/elconnectivities.m:147-212).  Returns (npf*nfe, ne)."""
    npf, nfe = perm.shape
    nd = dgnodes.shape[1] if dgnodes.shape[1] <= 3 else 3
    ne = dgnodes.shape[2]
    X = dgnodes[:, :nd, :]
    fn = np.transpose(X[perm, :, :], (0, 1, 3, 2))      # (npf, nfe, ne, nd)
    elcon = np.empty((npf, nfe, ne), dtype=np.int64)
    e1 = f[:, -2]
    e2 = f[:, -1]
    nf = f.shape[0]
    j1 = np.argmax(t2f[e1] == np.arange(nf)[:, None], axis=1)
    ref = fn[:, j1, e1, :]                               # (npf, nf, nd)
    ids = np.arange(nf)[None, :] * npf + np.arange(npf)[:, None]
    elcon[:, j1, e1] = ids
    inner = np.nonzero(e2 >= 0)[0]
    if inner.size:
        ee = e2[inner]
        j2 = np.argmax(t2f[ee] == inner[:, None], axis=1)
        other = fn[:, j2, ee, :]                          # (npf, nin, nd)
        d2 = ((other[:, None, :, :] - ref[None, :, inner, :]) ** 2).sum(-1)  # (k2, k1, nin)
        match = np.argmin(d2, axis=1)                    # for each node of side 2, node of side 1
        if np.max(np.min(d2, axis=1)) > 1e-16:
            raise ValueError("trace nodes do not match across a face")
        elcon[:, j2, ee] = ids[match, inner[None, :]]
    return np.asfortranarray(elcon.reshape(npf * nfe, ne, order="F")), nf * npf


def mkmesh(p, t, master, ncd_extra=0):
    nd = p.shape[1]
    f, t2f = mkt2f(t, nd)
    f = setbndnbrs(p, f, nd)
    bf_f = f[:, -1].copy()
    # bf = f(abs(t2f'),end): the face's last column (neighbour elt or -id)
    bf = np.where(bf_f[t2f] < 0, bf_f[t2f], bf_f[t2f] + 1).T
    dg = createnodes(p, t, master.plocvl)
    if ncd_extra:
        dg = np.concatenate([dg, np.zeros((dg.shape[0], ncd_extra, dg.shape[2]))], axis=1)
    elcon, nsiz = hdgelcon(dg, master.perm, f, t2f)
    return Mesh(nd=nd, porder=master.porder, p=p, t=t, f=f, t2f=t2f,
                bf=np.asfortranarray(bf.astype(np.int32)), dgnodes=np.asfortranarray(dg),
                elcon=elcon, nsiz=nsiz)


def mkmesh_square(nx, ny, master, jitter_amp=0.0, seed=0, ncd_extra=0):
    """Jittered unit square, 2*(nx)*(ny) triangles (nx, ny cells per side)."""
    p, t = squaremesh(nx + 1, ny + 1)
    if jitter_amp:
        p = jitter(p, 1.0 / max(nx, ny), jitter_amp, seed)
    return mkmesh(p, t, master, ncd_extra)


def mkmesh_cube(n, master, jitter_amp=0.0, seed=0):
    """Jittered unit cube, 6*n^3 tetrahedra."""
    p, t = cubemesh(n + 1, n + 1, n + 1)
    if jitter_amp:
        p = jitter(p, 1.0 / n, jitter_amp, seed)
        t = _orient(p, t)
    return mkmesh(p, t, master)

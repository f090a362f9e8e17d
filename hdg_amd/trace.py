"""Trace system of the HDG Newton step on the device: host-side face
connectivity (built once per mesh) and the operators of include/hdgx.h's
trace section.

Mirrors the reference's dense-block trace path (``app.denseblock = 1``):
``mkf2f`` (mesh/mkmesh/mkf2f.m), ``mkDenseBlockSystem``
(kernel/mkDenseBlockSystem.m:35-95), ``gmres`` (kernel/gmres.m) and, for the
block ILU(0) variant, ``computeBILU0`` / ``applyBILU0`` / ``gmres_bilu0``
(kernel/computeBILU0.m, applyBILU0.m, gmres_bilu0.m).  The reference orders the
ILU(0) rows by minimum discarded fill (kernel/computeMDFordering.m), a
sequential heuristic; here the rows are ordered by a greedy colouring of the
face graph so that every colour is one parallel sweep.
"""
import ctypes

import numpy as np

from .runtime import Dev, HdgxError, lib


class CTrace(ctypes.Structure):
    _fields_ = [("ncu", ctypes.c_int32), ("npf", ctypes.c_int32), ("nfe", ctypes.c_int32),
                ("ne", ctypes.c_int64), ("nf", ctypes.c_int64),
                ("fconn", ctypes.c_void_p), ("f2f", ctypes.c_void_p), ("elcon", ctypes.c_void_p),
                ("ncolors", ctypes.c_int32), ("color_ptr", ctypes.c_void_p),
                ("color_rows", ctypes.c_void_p), ("face_rank", ctypes.c_void_p)]


def face_connectivity(mesh):
    """fconn (4, nf): first element, its local face index, second element (-1
    on the boundary), its local face index -- the columns of mesh.f used by
    mkDenseBlockSystem.m:41-46 with the local indices of t2f."""
    f, t2f = mesh.f, mesh.t2f
    nf = f.shape[0]
    e1 = f[:, -2].astype(np.int64)
    e2 = f[:, -1].astype(np.int64)
    e2 = np.where(e2 >= 0, e2, -1)
    fid = np.arange(nf)
    j1 = np.argmax(t2f[e1] == fid[:, None], axis=1)
    j2 = np.where(e2 >= 0, np.argmax(t2f[np.maximum(e2, 0)] == fid[:, None], axis=1), 0)
    return np.asfortranarray(np.stack([e1, j1, e2, j2]).astype(np.int32))


def mkf2f(mesh):
    """(nbf, nf) face-to-face table, slot 0 the face, then the other faces of
    its first element and of its second, ascending local index; -1 where none
    (mesh/mkmesh/mkf2f.m:8-55, 0-based)."""
    fconn = face_connectivity(mesh)
    t2f = mesh.t2f
    nfe = t2f.shape[1]
    nbf = 2 * nfe - 1
    nf = fconn.shape[1]
    e1, j1, e2, j2 = (fconn[k].astype(np.int64) for k in range(4))
    f2f = -np.ones((nbf, nf), dtype=np.int32)
    f2f[0] = np.arange(nf)
    for k in range(1, nfe):
        lj = np.where(k - 1 < j1, k - 1, k)
        f2f[k] = t2f[e1, lj]
        lj2 = np.where(k - 1 < j2, k - 1, k)
        f2f[nfe - 1 + k] = np.where(e2 >= 0, t2f[np.maximum(e2, 0), lj2], -1)
    return np.asfortranarray(f2f)


def color_faces(f2f):
    """Greedy colouring of the face graph (faces adjacent when they share an
    element).  Returns (color_ptr, color_rows, face_rank, face_color): faces of
    colour c are color_rows[color_ptr[c]:color_ptr[c+1]] in ascending index,
    and face_rank is each face's position in that elimination order."""
    nbf, nf = f2f.shape
    f2f32 = np.asfortranarray(np.asarray(f2f, dtype=np.int32))
    color = np.empty(nf, dtype=np.int32)
    nc = ctypes.c_int32(0)
    rc = lib().hdgx_color_faces(nbf, nf, f2f32.ctypes.data_as(ctypes.c_void_p), color.ctypes.data_as(ctypes.c_void_p),
                                ctypes.byref(nc))
    if rc:
        raise HdgxError("hdgx_color_faces: bad face table")
    ncolors = nc.value
    order = np.lexsort((np.arange(nf), color))
    rank = np.empty(nf, dtype=np.int64)
    rank[order] = np.arange(nf)
    counts = np.bincount(color, minlength=ncolors)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return ptr, order.astype(np.int32), rank.astype(np.int32), color.astype(np.int32)


def check_coloring(f2f, color):
    nbr = f2f[1:]
    ok = nbr >= 0
    same = color[np.where(ok, nbr, 0)] == color[None, :]
    return not np.any(same & ok)


class TraceSystem:
    """Device-resident connectivity of one mesh plus the trace operators.

    All vectors are ``Dev`` arrays of shape (ncf, nf) (ncf = ncu*npf): the
    layout of UH (ncu, npf*nf)."""

    def __init__(self, ctx, mesh, ncu, with_bilu0=True, conn=None):
        """conn: optional host connectivity already built for this mesh
        (dict with fconn, f2f and color = color_faces(f2f))."""
        self.ctx = ctx
        self.mesh = mesh
        npf = mesh.elcon.shape[0] // mesh.t2f.shape[1]
        self.ncu, self.npf, self.nfe = ncu, npf, mesh.t2f.shape[1]
        self.ne, self.nf = mesh.t.shape[0], mesh.f.shape[0]
        self.ncf = ncu * npf
        self.nbf = 2 * self.nfe - 1
        conn = conn or {}
        self.fconn = Dev.from_numpy(conn["fconn"] if "fconn" in conn else face_connectivity(mesh), np.int32)
        self.f2f_host = conn["f2f"] if "f2f" in conn else mkf2f(mesh)
        self.f2f = Dev.from_numpy(self.f2f_host, np.int32)
        self.elcon = Dev.from_numpy(mesh.elcon, np.int32)
        c = CTrace(ncu, npf, self.nfe, self.ne, self.nf, self.fconn.ptr.value, self.f2f.ptr.value,
                   self.elcon.ptr.value, 0, None, None, None)
        if with_bilu0:
            if "color" in conn:
                ptr, rows, rank, color = conn["color"]
            else:
                ptr, rows, rank, color = color_faces(self.f2f_host)
                if not check_coloring(self.f2f_host, color):
                    raise HdgxError("face colouring is not proper")
            self.ncolors = len(ptr) - 1
            self._color_ptr = (ctypes.c_int32 * len(ptr))(*ptr.tolist())
            self.color_rows = Dev.from_numpy(rows, np.int32)
            self.face_rank = Dev.from_numpy(rank, np.int32)
            self.face_color = color
            c.ncolors = self.ncolors
            c.color_ptr = ctypes.cast(self._color_ptr, ctypes.c_void_p)
            c.color_rows = self.color_rows.ptr.value
            c.face_rank = self.face_rank.ptr.value
        self.c = c

    @classmethod
    def from_device(cls, ctx, ncu, npf, nfe, ne, nf, fconn, f2f, f2f_host, elcon, with_bilu0=True):
        """The same from connectivity already on the device (hdg_amd.devmesh):
        fconn (4, nf), f2f (2nfe-1, nf), elcon (npf*nfe, ne) Dev int32 arrays and
        f2f_host, the face table on the host for the face colouring."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.mesh = None
        self.ncu, self.npf, self.nfe = ncu, npf, nfe
        self.ne, self.nf = ne, nf
        self.ncf = ncu * npf
        self.nbf = 2 * nfe - 1
        self.fconn, self.f2f, self.elcon = fconn, f2f, elcon
        self.f2f_host = f2f_host
        c = CTrace(ncu, npf, nfe, ne, nf, fconn.ptr.value, f2f.ptr.value, elcon.ptr.value, 0, None, None, None)
        if with_bilu0:
            ptr, rows, rank, color = color_faces(f2f_host)
            self.ncolors = len(ptr) - 1
            self._color_ptr = (ctypes.c_int32 * len(ptr))(*ptr.tolist())
            self.color_rows = Dev.from_numpy(rows, np.int32)
            self.face_rank = Dev.from_numpy(rank, np.int32)
            self.face_color = color
            c.ncolors = self.ncolors
            c.color_ptr = ctypes.cast(self._color_ptr, ctypes.c_void_p)
            c.color_rows = self.color_rows.ptr.value
            c.face_rank = self.face_rank.ptr.value
        self.c = c
        return self

    # ---------------------------------------------------------------- operators
    def _check(self, rc):
        self.ctx._check(rc)

    def assemble(self, AE, FE, out=None):
        """(K, F, status) = mkDenseBlockSystem(AE, FE, ...), block rows scaled by
        the inverse diagonal block."""
        ncf, nbf, nf = self.ncf, self.nbf, self.nf
        if out is None:
            out = (Dev.empty((ncf, ncf, nbf - 1, nf)), Dev.empty((ncf, nf)), Dev.empty((nf,), "int32"))
        K, F, st = out
        self._check(lib().hdgx_trace_assemble(self.ctx.h, ctypes.byref(self.c), AE.ptr, FE.ptr, K.ptr, F.ptr, st.ptr))
        return K, F, st

    def matvec(self, K, x, y=None):
        y = y or Dev.empty((self.ncf, self.nf))
        self._check(lib().hdgx_trace_matvec(self.ctx.h, ctypes.byref(self.c), K.ptr, x.ptr, y.ptr))
        return y

    def bilu0_factor(self, K, LU=None):
        LU = LU or Dev.empty((self.ncf, self.ncf, self.nbf, self.nf))
        st = Dev.empty((self.nf,), "int32")
        self._check(lib().hdgx_bilu0_factor(self.ctx.h, ctypes.byref(self.c), K.ptr, LU.ptr, st.ptr))
        return LU, st

    def bilu0_apply(self, LU, x, y=None):
        y = y or Dev.empty((self.ncf, self.nf))
        self._check(lib().hdgx_bilu0_apply(self.ctx.h, ctypes.byref(self.c), LU.ptr, x.ptr, y.ptr))
        return y

    def gmres(self, K, F, x=None, LU=None, restart=20, tol=1e-6, maxit=0):
        """x = gmres(K, F, f2f, x, restart, tol, maxit) (kernel/gmres.m); with LU
        the block-ILU(0) left-preconditioned form (kernel/gmres_bilu0.m).
        Returns (x, iters, relres, flag)."""
        if x is None:
            x = Dev.zeros((self.ncf, self.nf))
        it, rr, fl = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        self._check(lib().hdgx_gmres(self.ctx.h, ctypes.byref(self.c), K.ptr, LU.ptr if LU is not None else None,
                                     F.ptr, x.ptr, restart, tol, maxit, ctypes.byref(it), ctypes.byref(rr),
                                     ctypes.byref(fl)))
        return x, it.value, rr.value, fl.value

    def gather(self, UH, UHE=None):
        ndf = self.npf * self.nfe
        UHE = UHE or Dev.empty((self.ncu, ndf, self.ne))
        self._check(lib().hdgx_trace_gather(self.ctx.h, self.ncu, self.npf, self.nfe, self.ne, self.elcon.ptr, UH.ptr,
                                            UHE.ptr))
        return UHE

    def scatter(self, Rhe, RH=None):
        """RH (ncf, nf) = the per-element Rh (ncu, npf*nfe, ne) summed at the
        global trace unknowns (kernel/hdg_residual.m:41-52)."""
        RH = RH or Dev.empty((self.ncf, self.nf))
        self._check(lib().hdgx_trace_scatter(self.ctx.h, ctypes.byref(self.c), Rhe.ptr, RH.ptr))
        return RH

    def local_update(self, npv, nco, DUDG, DUDG_DUH, DUH, DUDGT=None):
        DUDGT = DUDGT or Dev.empty((npv, nco, self.ne))
        self._check(lib().hdgx_local_update(self.ctx.h, npv, nco, self.ncu, self.npf, self.nfe, self.ne,
                                            self.elcon.ptr, DUDG.ptr, DUDG_DUH.ptr, DUH.ptr, DUDGT.ptr))
        return DUDGT

    def newton_update(self, npv, nc, nco, alpha, UDG0, DUDGT, UH0, DUH, UDG, UH):
        nuh = self.ncf * self.nf
        self._check(lib().hdgx_newton_update(self.ctx.h, npv, nc, nco, self.ne, nuh, alpha, UDG0.ptr, DUDGT.ptr,
                                             UH0.ptr, DUH.ptr, UDG.ptr, UH.ptr))

    def norm2(self, x):
        r = ctypes.c_double()
        n = int(np.prod(x.shape))
        self._check(lib().hdgx_norm2(self.ctx.h, n, x.ptr, ctypes.byref(r)))
        return r.value

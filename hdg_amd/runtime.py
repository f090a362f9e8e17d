"""Host runtime over the hdgx C ABI (include/hdgx.h), with torch for device
memory and streams.

The product path runs only on the HIP library: if libhdgx.so is missing or
no GPU is visible, every call raises -- there is no CPU fallback.

Functions mirror the reference's operator interface:
``hdg_elemental`` (kernel/hdg_elemental.m:1), ``schur_u`` / ``schur_uq`` (the MEX
files kernel/schur_u.c, kernel/schur_uq.c) and ``elemental_matrices`` (the
volint + faceint outputs).  Arrays keep the reference's MATLAB shapes; on the
device they are flat column-major torch tensors (``Dev``).
"""
import ctypes
import os

import numpy as np

from .app import App, CApp, MasterC

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhdgx.so")

PREC_FP64 = 0
PREC_FP32_IR = 1
INPLACE = 1          # HDGX_INPLACE

_lib = None


class HdgxError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise HdgxError(f"{LIB} is not built; run python -m hdg_amd.build")
        L = ctypes.CDLL(LIB)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.hdgx_create.argtypes = [ctypes.POINTER(vp), i32]
        L.hdgx_destroy.argtypes = [vp]
        L.hdgx_set_stream.argtypes = [vp, vp]
        L.hdgx_set_precision.argtypes = [vp, i32]
        L.hdgx_set_flags.argtypes = [vp, i32]
        L.hdgx_set_workspace_limit.argtypes = [vp, i64]
        L.hdgx_set_timing.argtypes = [vp, i32]
        L.hdgx_kernel_times.argtypes = [vp, vp, vp, vp]
        L.hdgx_last_error.argtypes = [vp]
        L.hdgx_last_error.restype = ctypes.c_char_p
        L.hdgx_elemental.argtypes = [vp, vp, vp, i64] + [vp] * 10
        L.hdgx_elemental_noalloc.argtypes = [vp, vp, vp, i64] + [vp] * 10
        L.hdgx_elemental_matrices.argtypes = [vp, vp, vp, i64] + [vp] * 15
        L.hdgx_residual.argtypes = [vp, vp, vp, i64] + [vp] * 8
        L.hdgx_getq.argtypes = [vp, vp, i32, i32, i32, i64] + [vp] * 4 + [ctypes.c_double, vp, i64, vp, vp]
        L.hdgx_schur_u.argtypes = [vp, i32, i32, i32, i32, i64] + [vp] * 11
        L.hdgx_schur_uq.argtypes = [vp, i32, i32, i32, i32, i32, i64] + [vp] * 15
        L.hdgx_elemental_batch.argtypes = [vp, vp, i32, vp]
        L.hdgx_elemental_batch_mixed.argtypes = [vp, i32, vp, vp]
        L.hdgx_plan_create.argtypes = [vp, i32, vp, vp, i32, i32, ctypes.POINTER(vp)]
        L.hdgx_plan_launch.argtypes = [vp, vp]
        L.hdgx_plan_capture.argtypes = [vp, vp]
        L.hdgx_plan_destroy.argtypes = [vp, vp]
        L.hdgx_trace_assemble.argtypes = [vp] * 7
        L.hdgx_trace_matvec.argtypes = [vp] * 5
        L.hdgx_bilu0_factor.argtypes = [vp] * 5
        L.hdgx_bilu0_apply.argtypes = [vp] * 5
        L.hdgx_gmres.argtypes = [vp] * 6 + [i32, ctypes.c_double, i32, vp, vp, vp]
        L.hdgx_trace_gather.argtypes = [vp, i32, i32, i32, i64, vp, vp, vp]
        L.hdgx_trace_scatter.argtypes = [vp, vp, vp, vp]
        L.hdgx_color_faces.argtypes = [i32, i64, vp, vp, vp]
        L.hdgx_local_update.argtypes = [vp, i32, i32, i32, i32, i32, i64, vp, vp, vp, vp, vp]
        L.hdgx_local_solve.argtypes = [vp, vp, vp, i64] + [vp] * 9
        L.hdgx_newton_update.argtypes = [vp, i32, i32, i32, i64, i64, ctypes.c_double, vp, vp, vp, vp, vp, vp]
        L.hdgx_norm2.argtypes = [vp, i64, vp, vp]
        L.hdgx_mesh_faces.argtypes = [vp, i32, i64, i32, vp, vp, vp, vp]
        L.hdgx_mesh_elcon.argtypes = [vp, i32, i32, i32, i32, i64, i64, vp, vp, vp, vp, vp]
        L.hdgx_mesh_f2f.argtypes = [vp, i32, i64, i64, vp, vp, vp, vp]
        L.hdgx_mesh_block_crs.argtypes = [vp, i32, i64, vp, vp, vp, vp]
        _lib = L
    return _lib


EXPORTS = ["hdgx_create", "hdgx_destroy", "hdgx_set_stream", "hdgx_set_precision", "hdgx_set_flags",
           "hdgx_set_workspace_limit", "hdgx_set_timing", "hdgx_kernel_times", "hdgx_last_error",
           "hdgx_elemental", "hdgx_elemental_matrices", "hdgx_residual", "hdgx_getq", "hdgx_schur_u", "hdgx_schur_uq",
           "hdgx_elemental_batch", "hdgx_elemental_batch_mixed", "hdgx_elemental_noalloc",
           "hdgx_plan_create", "hdgx_plan_launch", "hdgx_plan_capture", "hdgx_plan_destroy",
           "hdgx_trace_assemble", "hdgx_trace_matvec", "hdgx_bilu0_factor", "hdgx_bilu0_apply", "hdgx_gmres",
           "hdgx_trace_gather", "hdgx_trace_scatter", "hdgx_color_faces", "hdgx_local_update", "hdgx_local_solve", "hdgx_newton_update", "hdgx_norm2",
           "hdgx_mesh_faces", "hdgx_mesh_elcon", "hdgx_mesh_f2f", "hdgx_mesh_block_crs"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise HdgxError("no GPU visible: the hdgx path runs only on the device")
    return torch


class Dev:
    """A device array: flat column-major float64/int32 torch tensor + MATLAB shape."""

    def __init__(self, t, shape):
        self.t = t
        self.shape = tuple(int(s) for s in shape)

    @property
    def ptr(self):
        return ctypes.c_void_p(self.t.data_ptr())

    def numpy(self):
        return self.t.detach().cpu().numpy().reshape(self.shape, order="F")

    @staticmethod
    def from_numpy(a, dtype=None, device="cuda"):
        torch = _torch()
        a = np.asarray(a)
        if dtype is not None:
            a = a.astype(dtype)
        flat = np.ascontiguousarray(a.ravel(order="F"))
        return Dev(torch.from_numpy(flat).to(device), a.shape)

    @staticmethod
    def empty(shape, dtype="float64", device="cuda"):
        torch = _torch()
        n = int(np.prod(shape)) if len(shape) else 1
        td = torch.float64 if dtype == "float64" else torch.int32
        return Dev(torch.empty(max(n, 1), dtype=td, device=device), shape)

    @staticmethod
    def zeros(shape, dtype="float64", device="cuda"):
        d = Dev.empty(shape, dtype, device)
        d.t.zero_()
        return d


def _dev(a, dtype="float64"):
    if a is None:
        return None
    if isinstance(a, Dev):
        return a
    return Dev.from_numpy(a, np.float64 if dtype == "float64" else np.int32)


def _p(a):
    return None if a is None else a.ptr


class CSegment(ctypes.Structure):
    _fields_ = [("ne", ctypes.c_int64)] + [(n, ctypes.c_void_p) for n in (
        "dgnodes", "bf", "udg", "uh", "sh", "ae", "fe", "dudg", "dudg_duh", "status", "app")]


class Context:
    """One hdgx context (per host thread / stream) on a device."""

    def __init__(self, device=0, precision=PREC_FP64):
        torch = _torch()
        self.device = device
        h = ctypes.c_void_p()
        rc = lib().hdgx_create(ctypes.byref(h), device)
        if rc:
            raise HdgxError(f"hdgx_create failed ({rc})")
        self.h = h
        self._masters = {}
        self.set_precision(precision)
        self.bind_stream(torch.cuda.current_stream(device))

    def bind_stream(self, stream):
        self.stream = stream
        lib().hdgx_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream))

    def set_precision(self, precision):
        self.precision = precision
        self._check(lib().hdgx_set_precision(self.h, precision))

    def set_flags(self, flags):
        """hdgx_set_flags: INPLACE makes schur_u / schur_uq overwrite their
        (device) inputs as the MEX files do."""
        self._check(lib().hdgx_set_flags(self.h, int(flags)))

    def set_workspace_limit(self, nbytes):
        """hdgx_set_workspace_limit: most bytes of the element-matrix
        workspace (0 = the 24 GiB default); larger calls run in chunks."""
        self._check(lib().hdgx_set_workspace_limit(self.h, int(nbytes)))

    def set_timing(self, on):
        """hdgx_set_timing: events around the elemental kernels (synchronous calls)."""
        self._check(lib().hdgx_set_timing(self.h, int(bool(on))))

    def kernel_times(self):
        """(quadrature ms, elimination ms, launch pairs) summed since set_timing."""
        q, el, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._check(lib().hdgx_kernel_times(self.h, ctypes.byref(q), ctypes.byref(el), ctypes.byref(n)))
        return q.value, el.value, n.value

    def _check(self, rc):
        if rc:
            msg = lib().hdgx_last_error(self.h).decode()
            raise HdgxError(f"hdgx error {rc}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            lib().hdgx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def master_c(self, master):
        key = id(master)
        if key not in self._masters:
            self._masters[key] = (master, MasterC(master))
        return self._masters[key][1]

    # ------------------------------------------------------------ operators
    def elemental(self, master, app, dgnodes, bf, udg, uh, sh=None, out=None):
        """[AE, FE, DUDG, DUDG_DUH, status] = hdg_elemental(master, app, dgnodes, bf, UDG, UH, SH)."""
        ca = app.to_c() if isinstance(app, App) else app
        mc = self.master_c(master)
        dg, b, u, h = _dev(dgnodes), _dev(bf, "int32"), _dev(udg), _dev(uh)
        s = _dev(sh)
        ne = dg.shape[2]
        npv, ncu = master.npv, ca.ncu
        nco = ca.nc if ca.getdqdg == 1 else (ncu if ca.getdqdg == 0 else 0)   # hdg_elemental.m:20-26
        ndf = master.npf * master.nfe
        if out is None:
            out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nco, ne)),
                   Dev.empty((npv, nco, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
        ae, fe, dudg, dudg_duh, st = out
        self._check(lib().hdgx_elemental(self.h, ctypes.byref(mc.c), ctypes.byref(ca), ne, dg.ptr, b.ptr, u.ptr,
                                         h.ptr, _p(s), ae.ptr, fe.ptr, _p(dudg), _p(dudg_duh), st.ptr))
        return ae, fe, dudg, dudg_duh, st

    def local_solve(self, master, app, dgnodes, bf, udg, uh, sh, elcon, DUH, out=None, status=None):
        """hdgx_local_solve: DUDGT (npv, nc, ne) = dudg + dudg_duh DUH(:, elcon)
        recomputed element by element (kernel/hdg_local.m); uh the gathered
        trace (ncu, npf*nfe, ne), elcon (npf*nfe, ne) 0-based, DUH (ncu, nsiz)."""
        ca = app.to_c() if isinstance(app, App) else app
        mc = self.master_c(master)
        dg, b, u, h, s = _dev(dgnodes), _dev(bf, "int32"), _dev(udg), _dev(uh), _dev(sh)
        ec, dh = _dev(elcon, "int32"), _dev(DUH)
        ne = dg.shape[2]
        out = out or Dev.empty((master.npv, ca.nc, ne))
        self._check(lib().hdgx_local_solve(self.h, ctypes.byref(mc.c), ctypes.byref(ca), ne, dg.ptr, b.ptr, u.ptr,
                                           h.ptr, _p(s), ec.ptr, dh.ptr, out.ptr, _p(status)))
        return out

    def elemental_noalloc(self, master, app, dgnodes, bf, udg, uh, sh, out):
        """hdgx_elemental_noalloc: elemental() with every array a device array
        the caller already holds (`out` included); the library allocates
        nothing (its tables and workspace live in a fixed device array)."""
        ca = app.to_c() if isinstance(app, App) else app
        mc = self.master_c(master)
        dg, b, u, h, s = dgnodes, bf, udg, uh, sh
        ae, fe, dudg, dudg_duh, st = out
        self._check(lib().hdgx_elemental_noalloc(self.h, ctypes.byref(mc.c), ctypes.byref(ca), dg.shape[2], dg.ptr,
                                                 b.ptr, u.ptr, h.ptr, _p(s), ae.ptr, fe.ptr, _p(dudg), _p(dudg_duh),
                                                 _p(st)))
        return out

    def _segments(self, segments):
        keep = []
        arr = (CSegment * max(1, len(segments)))()
        for k, (app, dg, b, u, h, s, out) in enumerate(segments):
            ca = app.to_c() if isinstance(app, App) else app
            keep.append(ca)
            ae, fe, dudg, dudg_duh, st = out
            arr[k] = CSegment(dg.shape[2], dg.ptr.value, b.ptr.value, u.ptr.value, h.ptr.value,
                              s.ptr.value if s is not None else None, ae.ptr.value, fe.ptr.value, dudg.ptr.value,
                              dudg_duh.ptr.value, st.ptr.value if st is not None else None,
                              ctypes.cast(ctypes.pointer(ca), ctypes.c_void_p).value)
        return arr, keep

    def plan(self, masters, segments, mode=0, grid=0):
        """A reusable launch plan over segments of several orders (hdgx_plan_create):
        masters[k] is the master of segments[k]; mode 0 per order, 1 fused
        persistent, 2 per order on concurrent streams."""
        arr, keep = self._segments(segments)
        mcs = [self.master_c(m) for m in masters]
        mp = (ctypes.c_void_p * len(masters))(*[ctypes.cast(ctypes.pointer(mc.c), ctypes.c_void_p).value
                                                  for mc in mcs])
        h = ctypes.c_void_p()
        self._check(lib().hdgx_plan_create(self.h, len(segments), mp, arr, mode, grid, ctypes.byref(h)))
        return Plan(self, h, keep=(arr, keep, mcs, segments))

    def elemental_batch(self, master, segments):
        """One batched elemental launch over many requests' element ranges.
        segments: list of (app, dgnodes, bf, udg, uh, sh, (ae, fe, dudg, dudg_duh, status)) with
        Dev arrays; every request keeps its own physics parameters."""
        mc = self.master_c(master)
        keep = []
        arr = (CSegment * max(1, len(segments)))()
        for k, (app, dg, b, u, h, s, out) in enumerate(segments):
            ca = app.to_c() if isinstance(app, App) else app
            keep.append(ca)
            ae, fe, dudg, dudg_duh, st = out
            arr[k] = CSegment(dg.shape[2], dg.ptr.value, b.ptr.value, u.ptr.value, h.ptr.value,
                              s.ptr.value if s is not None else None, ae.ptr.value, fe.ptr.value, dudg.ptr.value,
                              dudg_duh.ptr.value, st.ptr.value if st is not None else None,
                              ctypes.cast(ctypes.pointer(ca), ctypes.c_void_p).value)
        self._check(lib().hdgx_elemental_batch(self.h, ctypes.byref(mc.c), len(segments), arr))

    def elemental_batch_mixed(self, masters, segments):
        """hdgx_elemental_batch_mixed: segments of several orders in one call,
        masters[k] the master of segments[k] (segments as elemental_batch)."""
        if not segments:
            return
        mcs = [self.master_c(m) for m in masters]
        mp = (ctypes.c_void_p * len(masters))(*[ctypes.cast(ctypes.pointer(mc.c), ctypes.c_void_p).value
                                                for mc in mcs])
        arr, keep = self._segments(segments)
        self._check(lib().hdgx_elemental_batch_mixed(self.h, len(segments), mp, arr))

    def elemental_matrices(self, master, app, dgnodes, bf, udg, uh, sh=None, out=None):
        """The element matrices as volint + faceint hand them to schur_primal
        (BD, F, GK, H, Ru, Rh and, with q, M, C, E, Rq); `out` (a dict of a
        previous call's arrays) is written in place."""
        ca = app.to_c() if isinstance(app, App) else app
        mc = self.master_c(master)
        dg, b, u, h = _dev(dgnodes), _dev(bf, "int32"), _dev(udg), _dev(uh)
        s = _dev(sh)
        ne = dg.shape[2]
        npv, ncu, nc, nd = master.npv, ca.ncu, ca.nc, master.nd
        ndf = master.npf * master.nfe
        shapes = {"BD": (npv, ncu, npv, nc, ne), "F": (npv, ncu, ncu, ndf, ne), "GK": (ncu, ndf, npv, nc, ne),
                  "H": (ncu, ndf, ncu, ndf, ne), "Ru": (npv, ncu, ne), "Rh": (ncu, ndf, ne)}
        if ca.localsolve:
            shapes.update(M=(npv, npv, ne), C=(npv, npv, nd, ne), E=(npv, ndf, nd, ne), Rq=(npv, ncu, nd, ne))
        if out is None:
            out = {k: Dev.empty(v) for k, v in shapes.items()}
        elif set(out) != set(shapes) or any(tuple(out[k].shape) != v for k, v in shapes.items()):
            raise HdgxError("elemental_matrices: out does not match these sizes")
        ptrs = [_p(out.get(k)) for k in ("BD", "F", "GK", "H", "Ru", "Rh", "M", "C", "E", "Rq")]
        self._check(lib().hdgx_elemental_matrices(self.h, ctypes.byref(mc.c), ctypes.byref(ca), ne, dg.ptr, b.ptr,
                                                  u.ptr, h.ptr, _p(s), *ptrs))
        return out

    def residual(self, master, app, dgnodes, bf, udg, uh, sh=None, out=None):
        """(Ru, Rh, Rq) of hdg_residual's element loop (kernel/hdg_residual.m:1-37):
        the residual without Jacobians; Rh per element with hdg_residual's sign
        (TraceSystem.scatter assembles it); Rq None when localsolve = 0."""
        ca = app.to_c() if isinstance(app, App) else app
        mc = self.master_c(master)
        dg, b, u, h = _dev(dgnodes), _dev(bf, "int32"), _dev(udg), _dev(uh)
        s = _dev(sh)
        ne = dg.shape[2]
        npv, ncu, nd = master.npv, ca.ncu, master.nd
        ndf = master.npf * master.nfe
        if out is None:
            out = (Dev.empty((npv, ncu, ne)), Dev.empty((ncu, ndf, ne)),
                   Dev.empty((npv, ncu, nd, ne)) if ca.localsolve else None)
        Ru, Rh, Rq = out
        self._check(lib().hdgx_residual(self.h, ctypes.byref(mc.c), ctypes.byref(ca), ne, dg.ptr, b.ptr, u.ptr, h.ptr,
                                        _p(s), Ru.ptr, Rh.ptr, _p(Rq)))
        return Ru, Rh, Rq

    def mesh_faces(self, t, nvert=None):
        """Faces on the device (mesh/mkmesh/mkt2f.m, 0-based): t (ne, nd+1)
        vertex ids.  Returns (f, t2f, nf) with f a Dev (nd+2, nf) -- column i
        is mesh.f's row i before boundary tagging -- and t2f a Dev (nd+1, ne),
        mesh.t2f transposed."""
        t = np.asarray(t)
        ne, nv = t.shape
        nd = nv - 1
        nvert = int(t.max()) + 1 if nvert is None else int(nvert)
        T = Dev.from_numpy(t.T, np.int32)
        t2f = Dev.empty((nv, ne), "int32")
        f = Dev.empty((nd + 2, ne * nv), "int32")
        nf = ctypes.c_int64(0)
        self._check(lib().hdgx_mesh_faces(self.h, nd, ne, nvert, T.ptr, t2f.ptr, f.ptr, ctypes.byref(nf)))
        n = nf.value
        return Dev(f.t[: (nd + 2) * max(n, 1)], (nd + 2, n)), t2f, n

    def mesh_elcon(self, dgnodes, perm, f, t2f, nf):
        """Trace-node numbering on the device (elconnectivities.m:147-212):
        f, t2f, nf as mesh_faces gives them.  Returns elcon, a Dev
        (npf*nfe, ne) equal to mesh.elcon."""
        dg = _dev(dgnodes)
        npv, ncd, ne = dg.shape
        perm = np.asarray(perm)
        npf, nfe = perm.shape
        P = Dev.from_numpy(perm, np.int32)
        out = Dev.empty((npf * nfe, ne), "int32")
        self._check(lib().hdgx_mesh_elcon(self.h, nfe - 1, npv, ncd, npf, ne, nf, dg.ptr, P.ptr, f.ptr, t2f.ptr,
                                          out.ptr))
        return out

    def mesh_f2f(self, f, t2f, nf):
        """Face-to-face table and face connectivity on the device (mkf2f.m):
        f, t2f, nf as mesh_faces gives them.  Returns (fconn (4, nf),
        f2f (2 nfe - 1, nf)) Dev arrays, equal to hdg_amd.trace's
        face_connectivity and mkf2f."""
        nd = f.shape[0] - 2
        ne = t2f.shape[1]
        fconn = Dev.empty((4, nf), "int32")
        f2f = Dev.empty((2 * nd + 1, nf), "int32")
        self._check(lib().hdgx_mesh_f2f(self.h, nd, ne, nf, f.ptr, t2f.ptr, fconn.ptr, f2f.ptr))
        return fconn, f2f

    def mesh_block_crs(self, f2f):
        """CBSR pattern of the face system on the device (block_crs.m with
        elcon2entcon.m on t2f'): f2f (2 nd + 1, nf) as mesh_f2f gives it.
        Returns (rowpts (nf+1,) from 0, colind (nblks,) 0-based) Dev arrays."""
        nbf, nf = f2f.shape
        nd = (nbf - 1) // 2
        rowpts = Dev.empty((nf + 1,), "int32")
        colind = Dev.empty((max(nf * nbf, 1),), "int32")
        nblks = ctypes.c_int64(0)
        self._check(lib().hdgx_mesh_block_crs(self.h, nd, nf, f2f.ptr, rowpts.ptr, colind.ptr, ctypes.byref(nblks)))
        return rowpts, Dev(colind.t[:nblks.value], (nblks.value,))

    def getq(self, master, dgnodes, udg, uh, sh=None, fc_q=1.0, ncu=None, out=None, inplace=False, shapnv=None):
        """QDG = getq(master, mesh, UDG, UH, SH, fc_q) (kernel/getq.m:1-129) for the
        element-gathered trace uh (ncu, npf*nfe, ne).  inplace=True writes it over
        UDG's q columns (hdg_solve.m:130-131; udg must then be a Dev) and returns
        udg; otherwise returns QDG (npv, ncu*nd, ne).  With shapnv =
        mkshape(p, plocvl, plocvl) it returns (QDG, qq)."""
        nd, npv = master.nd, master.npv
        ncu = ncu or nd + 2
        ncq = ncu * nd
        mc = self.master_c(master)
        if inplace and not isinstance(udg, Dev):
            raise HdgxError("getq(inplace=True) writes into udg: pass a device array")
        dg, u, h, s = _dev(dgnodes), _dev(udg), _dev(uh), _dev(sh)
        ne, nc, ncd = u.shape[2], u.shape[1], dg.shape[1]
        if inplace:
            if nc != ncu + ncq:
                raise HdgxError("getq(inplace=True) needs udg with nc = ncu*(nd+1) columns")
            qptr = ctypes.c_void_p(u.t.data_ptr() + 8 * npv * ncu)
            ldq, res = npv * nc, u
        else:
            res = out if out is not None else Dev.empty((npv, ncq, ne))
            qptr, ldq = res.ptr, npv * ncq
        sn = _dev(shapnv)
        qq = Dev.empty((npv, ncq, ne)) if sn is not None else None
        self._check(lib().hdgx_getq(self.h, ctypes.byref(mc.c), ncu, nc, ncd, ne, dg.ptr, u.ptr, h.ptr, _p(s),
                                    float(fc_q), qptr, ldq, _p(sn), _p(qq)))
        return (res, qq) if qq is not None else res

    def _schur_out(self, npv, ncu, nc, ndf, ne):
        return (Dev.empty((npv, nc, ne)), Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ncu, ndf, ncu, ndf, ne)),
                Dev.empty((ncu, ndf, ne)), Dev.empty((ne,), "int32"))

    def schur_u(self, BD, F, GK, H, Ru, Rh):
        """[DUDG, DUDG_DUH, AE, FE, status] = schur_u(BD, F, GK, H, Ru, Rh)."""
        BD, F, GK, H, Ru, Rh = (_dev(a) for a in (BD, F, GK, H, Ru, Rh))
        npv, ncu, _, nc, ne = BD.shape
        ndf = H.shape[1]
        out = self._schur_out(npv, ncu, nc, ndf, ne)
        self._check(lib().hdgx_schur_u(self.h, npv, ncu, nc, ndf, ne, BD.ptr, F.ptr, GK.ptr, H.ptr, Ru.ptr, Rh.ptr,
                                       *[o.ptr for o in out]))
        return out

    def schur_uq(self, M, C, E, BD, F, GK, H, Rq, Ru, Rh):
        """[DUDG, DUDG_DUH, AE, FE, status] = schur_uq(M, C, E, BD, F, GK, H, Rq, Ru, Rh)."""
        M, C, E, BD, F, GK, H, Rq, Ru, Rh = (_dev(a) for a in (M, C, E, BD, F, GK, H, Rq, Ru, Rh))
        npv, ne = M.shape[0], M.shape[2]
        ndf, nd = E.shape[1], E.shape[2]
        ncu, nc = BD.shape[1], BD.shape[3]
        out = self._schur_out(npv, ncu, nc, ndf, ne)
        self._check(lib().hdgx_schur_uq(self.h, nd, npv, ncu, nc, ndf, ne, M.ptr, C.ptr, E.ptr, BD.ptr, F.ptr,
                                        GK.ptr, H.ptr, Rq.ptr, Ru.ptr, Rh.ptr, *[o.ptr for o in out]))
        return out


class Plan:
    """Handle of an hdgx_plan: launch() enqueues the batch, capture() records
    it into a hipGraph that later launches replay."""

    def __init__(self, ctx, h, keep):
        self.ctx, self.h, self._keep = ctx, h, keep
        self.captured = False

    def launch(self):
        self.ctx._check(lib().hdgx_plan_launch(self.ctx.h, self.h))

    def capture(self):
        self.ctx._check(lib().hdgx_plan_capture(self.ctx.h, self.h))
        self.captured = True

    def close(self):
        if self.h:
            lib().hdgx_plan_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = {}


def context(device=0):
    if device not in _default:
        _default[device] = Context(device)
    return _default[device]


def hdg_elemental(master, app, dgnodes, bf, UDG, UH, SH=None, precision=PREC_FP64, device=0):
    """Host-array convenience form of the elemental operator: uploads, runs on
    the GPU and returns numpy arrays (AE, FE, DUDG, DUDG_DUH, status)."""
    ctx = context(device)
    ctx.set_precision(precision)
    outs = ctx.elemental(master, app, dgnodes, bf, UDG, UH, SH)
    ctx.stream.synchronize()
    return tuple(o.numpy() for o in outs)

"""ctypes binding of the oracle library (TEST INFRASTRUCTURE ONLY).

The oracle is the CPU FP64 restatement of the reference's elemental path
(oracle/hdg_oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []) + ["all"])


_native = None


def native_lib():
    """A -march=native build of the same source, for CPU timing only, built
    for the host it runs on into a temporary folder keyed by that host's CPU
    (a build made on another host is never loaded).  The process-wide
    library that every check uses stays the portable build: the two differ
    where an FP64 branch turns on rounding (|u.n| at u.n = 0 on a far-field
    face parallel to the flow), and checks are pinned on the portable one."""
    global _native
    if _native is None:
        import hashlib
        import platform
        import tempfile
        cpu = platform.processor() + platform.machine()
        try:
            with open("/proc/cpuinfo") as fh:
                cpu += "".join(l for l in fh if l.startswith(("model name", "flags")))
        except OSError:
            pass
        key = hashlib.sha1(cpu.encode()).hexdigest()[:12]
        out = os.path.join(tempfile.gettempdir(), f"hdgx_oracle_native_{key}", "liboracle_native.so")
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.check_call(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={out}"])
        _native = _bind(ctypes.CDLL(out))
    return _native


def _bind(L):
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    L.oracle_elemental.argtypes = [vp, vp, i64] + [vp] * 10 + [i32]
    L.oracle_elemental_matrices.argtypes = [vp, vp, i64] + [vp] * 15 + [i32]
    L.oracle_schur_u.argtypes = [i32] * 4 + [i64] + [vp] * 11 + [i32]
    L.oracle_schur_uq.argtypes = [i32] * 5 + [i64] + [vp] * 15 + [i32]
    L.oracle_schur_uq2.argtypes = [i32] * 5 + [i64] + [vp] * 15
    L.oracle_flux.argtypes = [vp, i32, vp, vp, vp, vp]
    L.oracle_fhat.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp]
    L.oracle_fbou.argtypes = [vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp]
    return L


def lib():
    """The portable build, which every check uses."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind(ctypes.CDLL(LIB))
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f(a):
    """Column-major contiguous float64 copy (memory order == MATLAB order)."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _ref(app):
    from hdg_amd.app import CApp
    return app if isinstance(app, CApp) else app.to_c()


def elemental(master, app, dgnodes, bf, udg, uh, sh=None, nthreads=0, native=False):
    """[AE, FE, DUDG, DUDG_DUH, status] = hdg_elemental(...) on the CPU
    (native=True: the -march=native build, for timing)."""
    from hdg_amd.app import MasterC
    mc = MasterC(master)
    ca = _ref(app)
    npv, ncu, nc = master.npv, ca.ncu, ca.nc
    ndf = master.npf * master.nfe
    ne = dgnodes.shape[2]
    dg, u, h = _f(dgnodes), _f(udg), _f(uh)
    s = None if sh is None else _f(sh)
    b = np.asfortranarray(np.asarray(bf, dtype=np.int32))
    ae = np.zeros((ncu, ndf, ncu, ndf, ne), order="F")
    fe = np.zeros((ncu, ndf, ne), order="F")
    dudg = np.zeros((npv, nc, ne), order="F")
    dudg_duh = np.zeros((npv, nc, ncu, ndf, ne), order="F")
    st = np.zeros(ne, dtype=np.int32)
    L = native_lib() if native else lib()
    rc = L.oracle_elemental(ctypes.byref(mc.c), ctypes.byref(ca), ne, _p(dg), _p(b), _p(u), _p(h), _p(s),
                                _p(ae), _p(fe), _p(dudg), _p(dudg_duh), _p(st), nthreads)
    if rc:
        raise RuntimeError(f"oracle_elemental failed ({rc})")
    return ae, fe, dudg, dudg_duh, st


def elemental_matrices(master, app, dgnodes, bf, udg, uh, sh=None, nthreads=0):
    """volint + faceint outputs: dict of BD, F, GK, H, Ru, Rh (+ M, C, E, Rq)."""
    from hdg_amd.app import MasterC
    mc = MasterC(master)
    ca = _ref(app)
    npv, ncu, nc, nd = master.npv, ca.ncu, ca.nc, master.nd
    ndf = master.npf * master.nfe
    ne = dgnodes.shape[2]
    out = {
        "BD": np.zeros((npv, ncu, npv, nc, ne), order="F"),
        "F": np.zeros((npv, ncu, ncu, ndf, ne), order="F"),
        "GK": np.zeros((ncu, ndf, npv, nc, ne), order="F"),
        "H": np.zeros((ncu, ndf, ncu, ndf, ne), order="F"),
        "Ru": np.zeros((npv, ncu, ne), order="F"),
        "Rh": np.zeros((ncu, ndf, ne), order="F"),
    }
    if ca.localsolve:
        out.update(M=np.zeros((npv, npv, ne), order="F"), C=np.zeros((npv, npv, nd, ne), order="F"),
                   E=np.zeros((npv, ndf, nd, ne), order="F"), Rq=np.zeros((npv, ncu, nd, ne), order="F"))
    dg, u, h = _f(dgnodes), _f(udg), _f(uh)
    s = None if sh is None else _f(sh)
    b = np.asfortranarray(np.asarray(bf, dtype=np.int32))
    rc = lib().oracle_elemental_matrices(
        ctypes.byref(mc.c), ctypes.byref(ca), ne, _p(dg), _p(b), _p(u), _p(h), _p(s),
        *[_p(out.get(k)) for k in ("BD", "F", "GK", "H", "Ru", "Rh", "M", "C", "E", "Rq")], nthreads)
    if rc:
        raise RuntimeError(f"oracle_elemental_matrices failed ({rc})")
    return out


def _schur_outputs(npv, ncu, nc, ndf, ne):
    return (np.zeros((npv, nc, ne), order="F"), np.zeros((npv, nc, ncu, ndf, ne), order="F"),
            np.zeros((ncu, ndf, ncu, ndf, ne), order="F"), np.zeros((ncu, ndf, ne), order="F"),
            np.zeros(ne, dtype=np.int32))


def schur_u(BD, F, GK, H, Ru, Rh, nthreads=0):
    """[DUDG, DUDG_DUH, AE, FE, status] = schur_u(BD,F,GK,H,Ru,Rh) (kernel/schur_u.c)."""
    npv, ncu, _, nc = BD.shape[:4]
    ne = BD.shape[4] if BD.ndim == 5 else 1
    ndf = H.shape[1]
    outs = _schur_outputs(npv, ncu, nc, ndf, ne)
    args = [_f(a) for a in (BD, F, GK, H, Ru, Rh)]
    lib().oracle_schur_u(npv, ncu, nc, ndf, ne, *[_p(a) for a in args], *[_p(o) for o in outs], nthreads)
    return outs


def schur_uq(M, C, E, BD, F, GK, H, Rq, Ru, Rh, nthreads=0, variant="chol"):
    """schur_uq MEX (kernel/schur_uq.c, variant 'chol') or schur_uq2 (inv(M) form,
    kernel/schur_primal.m:52-150, variant 'inv')."""
    npv = M.shape[0]
    ne = M.shape[2] if M.ndim == 3 else 1
    ndf, nd = E.shape[1], E.shape[2]
    ncu, nc = BD.shape[1], BD.shape[3]
    outs = _schur_outputs(npv, ncu, nc, ndf, ne)
    args = [_f(a) for a in (M, C, E, BD, F, GK, H, Rq, Ru, Rh)]
    if variant == "chol":
        lib().oracle_schur_uq(nd, npv, ncu, nc, ndf, ne, *[_p(a) for a in args], *[_p(o) for o in outs], nthreads)
    else:
        lib().oracle_schur_uq2(nd, npv, ncu, nc, ndf, ne, *[_p(a) for a in args], *[_p(o) for o in outs])
    return outs


def flux(app, pg, udg):
    ca = _ref(app)
    ng = pg.shape[0]
    f = np.zeros((ng, ca.ncu, ca.nd), order="F")
    fu = np.zeros((ng, ca.ncu, ca.nd, ca.nc), order="F")
    lib().oracle_flux(ctypes.byref(ca), ng, _p(_f(pg)), _p(_f(udg)), _p(f), _p(fu))
    return f, fu


def fhat(app, nl, pg, udg, uh):
    ca = _ref(app)
    ng = pg.shape[0]
    fh = np.zeros((ng, ca.ncu), order="F")
    fu = np.zeros((ng, ca.ncu, ca.nc), order="F")
    fh_uh = np.zeros((ng, ca.ncu, ca.ncu), order="F")
    lib().oracle_fhat(ctypes.byref(ca), ng, _p(_f(nl)), _p(_f(pg)), _p(_f(udg)), _p(_f(uh)), _p(fh), _p(fu), _p(fh_uh))
    return fh, fu, fh_uh


def fbou(app, ib, uinf, nl, pg, udg, uh):
    ca = _ref(app)
    ng = pg.shape[0]
    fh = np.zeros((ng, ca.ncu), order="F")
    fu = np.zeros((ng, ca.ncu, ca.nc), order="F")
    fh_uh = np.zeros((ng, ca.ncu, ca.ncu), order="F")
    rc = lib().oracle_fbou(ctypes.byref(ca), ib, _p(_f(uinf)), ng, _p(_f(nl)), _p(_f(pg)), _p(_f(udg)), _p(_f(uh)),
                           _p(fh), _p(fu), _p(fh_uh))
    if rc:
        raise RuntimeError(f"oracle_fbou unsupported boundary model {ib}")
    return fh, fu, fh_uh

"""The ``app`` struct of the elemental path and its C-ABI mirror.

Fields are the ones ``volint``/``faceint`` read (kernel/volint.m:10-23,
kernel/faceint.m:13-27): physics plugin, sizes, parameter vector, boundary
tables and the time/local-solve flags.  Physics plugins are chosen at compile
time on the device (Euler 2D/3D, Navier-Stokes 2D with optional laplacian
artificial viscosity), with the reference's parameter vectors:
Euler {gam, eps, tau} (problem/euler/euler_naca2.m:14-38, tau = param{end}),
Euler 3D {gam, eps, Re, Pr, Minf, tau} (fhat_euler3d.c reads tau = param[5]),
NS {gam, eps, Re, Pr, Minf, tau} (problem/ns/ns_naca.m:110-136).
"""
import ctypes
from dataclasses import dataclass, field
from typing import List

import numpy as np

PHYS_EULER = 0
PHYS_NS = 1
MAX_PARAM = 8
MAX_BC = 8

BC_FARFIELD = 1
BC_WALL = 2


class CMaster(ctypes.Structure):
    """hdgx_master (include/hdgx.h)."""
    _fields_ = [("nd", ctypes.c_int32), ("porder", ctypes.c_int32), ("npv", ctypes.c_int32),
                ("npf", ctypes.c_int32), ("nfe", ctypes.c_int32), ("ngv", ctypes.c_int32),
                ("ngf", ctypes.c_int32),
                ("shapvt", ctypes.c_void_p), ("shapvg", ctypes.c_void_p),
                ("shapvgdotshapvl", ctypes.c_void_p), ("shapft", ctypes.c_void_p),
                ("shapfg", ctypes.c_void_p), ("shapfgdotshapfc", ctypes.c_void_p),
                ("shapmv", ctypes.c_void_p), ("shapmf", ctypes.c_void_p),
                ("perm", ctypes.c_void_p), ("permgeom", ctypes.c_void_p)]


class CApp(ctypes.Structure):
    """hdgx_app (include/hdgx.h)."""
    _fields_ = [("physics", ctypes.c_int32), ("nd", ctypes.c_int32), ("ncu", ctypes.c_int32),
                ("nc", ctypes.c_int32), ("nch", ctypes.c_int32), ("ncd", ctypes.c_int32),
                ("localsolve", ctypes.c_int32), ("tdep", ctypes.c_int32), ("av_col", ctypes.c_int32),
                ("getdqdg", ctypes.c_int32),
                ("fc_u", ctypes.c_double), ("fc_q", ctypes.c_double), ("time", ctypes.c_double),
                ("nparam", ctypes.c_int32), ("param", ctypes.c_double * MAX_PARAM),
                ("nbc", ctypes.c_int32), ("bcm", ctypes.c_int32 * MAX_BC),
                ("bcs", ctypes.c_double * (MAX_BC * MAX_PARAM))]


@dataclass
class App:
    physics: int
    nd: int
    ncu: int
    nc: int
    param: List[float]
    bcm: List[int]
    bcs: np.ndarray                   # (nbc, ncu)
    localsolve: int = 0
    tdep: int = 0
    fc_u: float = 0.0
    fc_q: float = 1.0
    time: float = 0.0
    av_col: int = -1
    getdqdg: int = 1

    @property
    def nch(self):
        return self.ncu

    @property
    def ncd(self):
        return self.nd + (1 if self.av_col >= 0 else 0)

    def to_c(self):
        c = CApp()
        c.physics, c.nd, c.ncu, c.nc, c.nch, c.ncd = (self.physics, self.nd, self.ncu, self.nc,
                                                      self.nch, self.ncd)
        c.localsolve, c.tdep, c.av_col, c.getdqdg = self.localsolve, self.tdep, self.av_col, self.getdqdg
        c.fc_u, c.fc_q, c.time = self.fc_u, self.fc_q, self.time
        if len(self.param) > MAX_PARAM or len(self.bcm) > MAX_BC:
            raise ValueError("too many parameters or boundary ids")
        c.nparam = len(self.param)
        for i, v in enumerate(self.param):
            c.param[i] = float(v)
        c.nbc = len(self.bcm)
        bcs = np.asarray(self.bcs, dtype=np.float64).reshape(len(self.bcm), -1)
        for i, b in enumerate(self.bcm):
            c.bcm[i] = int(b)
            for k in range(bcs.shape[1]):
                c.bcs[i * MAX_PARAM + k] = bcs[i, k]
        return c


class MasterC:
    """Keeps the column-major copies alive while a CMaster points at them."""

    def __init__(self, master):
        m = master
        self._keep = {}

        def f64(name):
            a = np.ascontiguousarray(np.asarray(getattr(m, name), dtype=np.float64).ravel(order="F"))
            self._keep[name] = a
            return a.ctypes.data

        def i32(name):
            a = np.ascontiguousarray(np.asarray(getattr(m, name), dtype=np.int32).ravel(order="F"))
            self._keep[name] = a
            return a.ctypes.data

        self.c = CMaster(nd=m.nd, porder=m.porder, npv=m.npv, npf=m.npf, nfe=m.nfe, ngv=m.ngv,
                         ngf=m.ngf)
        for name in ("shapvt", "shapvg", "shapvgdotshapvl", "shapft", "shapfg", "shapfgdotshapfc",
                     "shapmv", "shapmf"):
            setattr(self.c, name, f64(name))
        self.c.perm = i32("perm")
        self.c.permgeom = i32("permgeom")


def euler_app(nd, gam=1.4, minf=0.5, tau=2.0, eps=0.0, alpha=0.0, nbc=None):
    """Steady Euler, all boundaries far field at the free stream
    ui = [1, cos a, sin a, 0.5 + p_inf/(gam-1)], p_inf = 1/(gam M^2)
    (problem/euler/euler_naca2.m:17-42)."""
    ncu = nd + 2
    pinf = 1.0 / (gam * minf ** 2)
    if nd == 2:
        ui = [1.0, np.cos(alpha), np.sin(alpha), 0.5 + pinf / (gam - 1.0)]
        param = [gam, eps, tau]
    else:
        ui = [1.0, np.cos(alpha), np.sin(alpha), 0.0, 0.5 + pinf / (gam - 1.0)]
        param = [gam, eps, 0.0, 0.0, minf, tau]
    nbc = nbc or (4 if nd == 2 else 6)
    return App(physics=PHYS_EULER, nd=nd, ncu=ncu, nc=ncu, param=param,
               bcm=[BC_FARFIELD] * nbc, bcs=np.tile(np.array(ui), (nbc, 1)))


def ns_app(gam=1.4, eps=0.1, re=1e4, pr=0.72, minf=0.3, tau=3.0, alpha=0.0, av=False):
    """Steady 2D Navier-Stokes with u and q (localsolve), far field all round
    (problem/ns/ns_naca.m:110-158); av=True reads a laplacian artificial
    viscosity from dgnodes(:,3,:)."""
    pinf = 1.0 / (gam * minf ** 2)
    ui = [1.0, np.cos(alpha), np.sin(alpha), 0.5 + pinf / (gam - 1.0)]
    return App(physics=PHYS_NS, nd=2, ncu=4, nc=12, param=[gam, eps, re, pr, minf, tau],
               bcm=[BC_FARFIELD] * 4, bcs=np.tile(np.array(ui), (4, 1)), localsolve=1,
               fc_q=1.0, av_col=2 if av else -1)

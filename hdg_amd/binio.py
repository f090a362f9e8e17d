"""The reference's binary problem files: read, write, and convert to and from
this package's Master / App / Mesh and state arrays.

The reference has two layouts:

* ``"matlab"``: what util/writeBinaryFile.m:14-96 writes for one process:
  the 100-entry ``ndims`` header, then the master tables, the app vectors
  (bcm, bcs, bcd, bcv, dt, param, flag, factor, problem), the mesh (dgnodes,
  elcon, bf, t2f, perm, permgeom), UDG, UH, and the trace system's CBSR
  pattern (rowpts, colind);
* ``"cpp"``: what kernel/readInputFile.cpp:84-196 reads and 211-327 writes:
  the same header, then UDG, UH, the mesh (dgnodes, elcon, bf, f, t2f, f2f,
  perm, permgeom), the master tables and the app vectors without ``problem``.

Every value is a float64 in native byte order; integer arrays are rounded on
read (readInputFile.cpp:73-82).  Arrays are column-major and hold MATLAB's
1-based indices; the app vectors follow util/preprocess.m:212-225:
flag = [tdep wave alag adjoint linearproblem flg_q flg_p flg_g],
factor = [fc_u fc_q fc_p dtfc alpha],
problem = [hybrid appname linearSolver jacobianStep orderingStep] with
appname 0 = Euler, 1 = Navier-Stokes.
"""
from dataclasses import dataclass, field
from typing import Dict

import numpy as np

from .app import App
from .master import Master
from .mesh import Mesh

NDIMS = 100
# header slots (0-based) of util/writeBinaryFile.m:16-51
HEADER = ["nd", "ncd", "nfe", "nve", "nvf", "ne", "nf", "nv", "ndh", "npe", "npf", "nme", "nmf", "nge", "ngf",
          "porder", "morder", "torder", "nstage", "nc", "ncu", "ncq", "ncp", "nch", "ns", "nb", "ndt", "nparam",
          "nflag", "nfactor", "nproblem", "cbsr_nrows", "cbsr_nblks", "maxBlocksPerRow", "blkSize"]
NPROC_SLOT = 39

_MASTER = [("plocvl", "d", lambda h: h["npe"] * h["nd"]),
           ("gpvl", "d", lambda h: h["nge"] * h["nd"]),
           ("gwvl", "d", lambda h: h["nge"]),
           ("plocfc", "d", lambda h: h["npf"] * (h["nd"] - 1)),
           ("gpfc", "d", lambda h: h["ngf"] * (h["nd"] - 1)),
           ("gwfc", "d", lambda h: h["ngf"]),
           ("shapvt", "d", lambda h: h["nge"] * h["npe"] * (h["nd"] + 1)),
           ("shapvg", "d", lambda h: h["npe"] * h["nge"] * (h["nd"] + 1)),
           ("shapvgdotshapvl", "d", lambda h: h["npe"] ** 2 * h["nge"] * (h["nd"] + 1)),
           ("shapft", "d", lambda h: h["ngf"] * h["npf"] * h["nd"]),
           ("shapfg", "d", lambda h: h["npf"] * h["ngf"] * h["nd"]),
           ("shapfgdotshapfc", "d", lambda h: h["npf"] ** 2 * h["ngf"] * h["nd"]),
           ("shapmv", "d", lambda h: h["nge"] * h["nme"] * (h["nd"] + 1)),
           ("shapmf", "d", lambda h: h["ngf"] * h["nmf"] * h["nd"])]
_APP = [("bcm", "i", lambda h: h["nb"]),
        ("bcs", "d", lambda h: h["nb"] * h["nch"]),
        ("bcd", "i", lambda h: h["nb"]),
        ("bcv", "d", lambda h: h["nb"] * h["nch"]),
        ("dt", "d", lambda h: h["ndt"]),
        ("param", "d", lambda h: h["nparam"]),
        ("flag", "i", lambda h: h["nflag"]),
        ("factor", "d", lambda h: h["nfactor"])]
_UDG = ("UDG", "d", lambda h: h["npe"] * h["nc"] * h["ne"])
_UH = ("UH", "d", lambda h: h["nch"] * h["ndh"])
_DG = ("dgnodes", "d", lambda h: h["nme"] * h["ncd"] * h["ne"])
_ELCON = ("elcon", "i", lambda h: h["npf"] * h["nfe"] * h["ne"])
_BF = ("bf", "i", lambda h: h["nfe"] * h["ne"])
_T2F = ("t2f", "i", lambda h: h["ne"] * h["nfe"])
_PERM = ("perm", "i", lambda h: h["npf"] * h["nfe"])
_PERMGEOM = ("permgeom", "i", lambda h: h["nmf"] * h["nfe"])

LAYOUTS = {
    # util/writeBinaryFile.m:52-92 (PDG, written only for app.wave = 1, is not supported)
    "matlab": _MASTER + _APP + [("problem", "i", lambda h: h["nproblem"]),
                                _DG, _ELCON, _BF, _T2F, _PERM, _PERMGEOM, _UDG, _UH,
                                ("cbsr_rowpts", "i", lambda h: h["cbsr_nrows"] + 1),
                                ("cbsr_colind", "i", lambda h: h["cbsr_nblks"])],
    # kernel/readInputFile.cpp:129-194
    "cpp": [_UDG, _UH, _DG, _ELCON, _BF,
            ("f", "i", lambda h: h["nf"] * (h["nvf"] + 2)),
            _T2F,
            ("f2f", "i", lambda h: h["nf"] * (2 * h["nfe"] - 1)),
            ("perm", "i", lambda h: h["nfe"] * h["npf"]),
            _PERMGEOM] + _MASTER + _APP,
}


@dataclass
class ProblemFile:
    """Header and arrays of one file, in the reference's conventions (flat
    column-major arrays, 1-based indices)."""
    header: Dict[str, int]
    arrays: Dict[str, np.ndarray] = field(default_factory=dict)
    layout: str = "matlab"


def _header_vector(h):
    v = np.zeros(NDIMS)
    for i, name in enumerate(HEADER):
        v[i] = h.get(name, 0)
    v[NPROC_SLOT] = h.get("nproc", 1)
    return v


def write_file(path, pf):
    """Writes pf in its layout; sizes must match the header."""
    h = pf.header
    with open(path, "wb") as fh:
        fh.write(_header_vector(h).astype("<f8" if np.little_endian else ">f8").tobytes())
        for name, kind, size in LAYOUTS[pf.layout]:
            a = np.asarray(pf.arrays[name], dtype=np.float64).ravel(order="F")
            if a.size != size(h):
                raise ValueError(f"{name}: {a.size} values, the header implies {size(h)}")
            fh.write(a.tobytes())


def read_file(path, layout="matlab"):
    """Reads a file of the given layout into a ProblemFile (integer arrays
    rounded to int64)."""
    raw = np.fromfile(path, dtype=np.float64)
    if raw.size < NDIMS:
        raise ValueError("file shorter than its header")
    hv = raw[:NDIMS]
    h = {name: int(round(hv[i])) for i, name in enumerate(HEADER)}
    h["nproc"] = int(round(hv[NPROC_SLOT]))
    if h["nproc"] > 1:
        raise ValueError("a per-subdomain file set (nproc > 1) is not supported")
    pos = NDIMS
    arrays = {}
    for name, kind, size in LAYOUTS[layout]:
        n = size(h)
        if pos + n > raw.size:
            raise ValueError(f"file ends inside {name}")
        a = raw[pos:pos + n]
        arrays[name] = np.rint(a).astype(np.int64) if kind == "i" else a.copy()
        pos += n
        if layout == "matlab" and name == "flag" and n > 1 and arrays[name][1] != 0:
            raise ValueError("wave problems (PDG block) are not supported")
    if pos != raw.size:
        raise ValueError(f"{raw.size - pos} values after the last array")
    return ProblemFile(h, arrays, layout)


# ---------------------------------------------------------------- conversion
def block_crs(t2f):
    """CBSR pattern of the face system (mesh/mkmesh/block_crs.m with
    elcon2entcon.m:27-45 on t2f'): row i lists face i, then the other faces
    sharing an element with it, ascending.  t2f (ne, nfe) 0-based; returns
    rowpts (nf+1, offsets from 0) and colind (1-based face ids, as MATLAB
    stores them)."""
    ne, nfe = t2f.shape
    nf = int(t2f.max()) + 1 if t2f.size else 0
    e_of = np.repeat(np.arange(ne), nfe)
    f_of = t2f.ravel()
    order = np.argsort(f_of, kind="stable")
    starts = np.searchsorted(f_of[order], np.arange(nf + 1))
    rowpts = [0]
    colind = []
    for i in range(nf):
        elems = e_of[order[starts[i]:starts[i + 1]]]
        k = np.unique(t2f[elems].ravel())
        colind.append(i)
        colind.extend(int(v) for v in k if v != i)
        rowpts.append(len(colind))
    return np.array(rowpts, dtype=np.int64), np.array(colind, dtype=np.int64) + 1


def faces_from_t2f(t2f, bf):
    """Element columns of mesh.f from t2f and bf (mesh/mkmesh/mkt2f.m's order):
    the first element visiting a face over (element, local face) and the
    second, or the boundary id -k taken from bf.  0-based; returns (nf, 2)."""
    ne, nfe = t2f.shape
    nf = int(t2f.max()) + 1 if t2f.size else 0
    ef = np.full((nf, 2), -1, dtype=np.int64)
    seen = np.zeros(nf, dtype=bool)
    for e in range(ne):
        for j in range(nfe):
            fi = t2f[e, j]
            if not seen[fi]:
                ef[fi, 0] = e
                ef[fi, 1] = bf[j, e] if bf[j, e] < 0 else -1
                seen[fi] = True
            else:
                ef[fi, 1] = e
    return ef


def export_problem(master, app, mesh, UDG, UH, dt=(1.0,), torder=1, nstage=1, layout="matlab", problem_extra=(0, 0, 0)):
    """A ProblemFile of this package's objects: 1-based indices, the app
    vectors of util/preprocess.m:212-225 (flags the elemental path does not
    use are written as 0), the CBSR pattern of the trace system."""
    nd, npv, npf, nfe = master.nd, master.npv, master.npf, master.nfe
    ne, nf = mesh.ne, mesh.nf
    ncu, nc = app.ncu, app.nc
    ncd = mesh.dgnodes.shape[1]
    t2f = np.asarray(mesh.t2f)
    rowpts, colind = block_crs(t2f)
    nb = len(app.bcm)
    h = dict(nd=nd, ncd=ncd, nfe=nfe, nve=nd + 1, nvf=nd, ne=ne, nf=nf, nv=mesh.p.shape[0], ndh=mesh.nsiz,
             npe=npv, npf=npf, nme=npv, nmf=npf, nge=master.ngv, ngf=master.ngf, porder=master.porder,
             morder=mesh.porder, torder=torder, nstage=nstage, nc=nc, ncu=ncu, ncq=nc - ncu, ncp=0, nch=app.nch,
             ns=1, nb=nb, ndt=len(dt), nparam=len(app.param), nflag=8, nfactor=5, nproblem=5,
             cbsr_nrows=nf, cbsr_nblks=int(colind.size), maxBlocksPerRow=int(np.diff(rowpts).max()) if nf else 0,
             blkSize=ncu * npf, nproc=1)
    f = np.asarray(mesh.f).copy()
    nvf = nd
    f[:, :nvf] += 1
    f[:, nvf] += 1
    f[:, nvf + 1] = np.where(f[:, nvf + 1] >= 0, f[:, nvf + 1] + 1, f[:, nvf + 1])
    from .trace import mkf2f
    f2f = np.asarray(mkf2f(mesh)).astype(np.int64) + 1          # 0 where there is none
    bcs = np.asarray(app.bcs, dtype=np.float64).reshape(nb, -1)
    A = {
        "plocvl": master.plocvl, "gpvl": master.gpvl, "gwvl": master.gwvl, "plocfc": master.plocfc,
        "gpfc": master.gpfc, "gwfc": master.gwfc, "shapvt": master.shapvt, "shapvg": master.shapvg,
        "shapvgdotshapvl": master.shapvgdotshapvl, "shapft": master.shapft, "shapfg": master.shapfg,
        "shapfgdotshapfc": master.shapfgdotshapfc, "shapmv": master.shapmv, "shapmf": master.shapmf,
        "bcm": np.asarray(app.bcm), "bcs": bcs.T, "bcd": np.zeros(nb), "bcv": np.zeros((ncu, nb)),
        "dt": np.asarray(dt, dtype=np.float64), "param": np.asarray(app.param, dtype=np.float64),
        "flag": np.array([app.tdep, 0, 0, 0, 0, app.localsolve, 0, 0]),
        "factor": np.array([app.fc_u, app.fc_q, 0.0, 0.0, 0.0]),
        "problem": np.array([problem_extra[0], app.physics, problem_extra[1], problem_extra[2], 0]),
        "dgnodes": mesh.dgnodes, "elcon": np.asarray(mesh.elcon) + 1, "bf": mesh.bf, "t2f": t2f + 1,
        "f": f, "f2f": f2f, "perm": np.asarray(master.perm) + 1, "permgeom": np.asarray(master.permgeom) + 1,
        "UDG": UDG, "UH": UH, "cbsr_rowpts": rowpts, "cbsr_colind": colind,
    }
    keep = {name for name, _, _ in LAYOUTS[layout]}
    return ProblemFile(h, {k: v for k, v in A.items() if k in keep}, layout)


def import_problem(pf, physics=None, av_col=None):
    """(master, app, mesh, UDG, UH) from a ProblemFile: the master tables as
    stored (the kernels take any consistent tables; pgauss is not stored and
    is taken as 2p, the built quadrature), the app of the header and app
    vectors (physics from problem(2) unless given), a mesh whose vertex
    lists are absent (p empty, t of -1: the elemental and trace paths use
    dgnodes, bf, t2f, elcon and the element columns of f)."""
    h, a = pf.header, pf.arrays
    nd, npe, npf, nfe, ne = h["nd"], h["npe"], h["npf"], h["nfe"], h["ne"]
    nge, ngf, nme, nmf = h["nge"], h["ngf"], h["nme"], h["nmf"]
    if nme != npe or nmf != npf:
        raise ValueError("only isoparametric problems (morder = porder) are supported")

    def F(name, *shape):
        return np.asfortranarray(np.asarray(a[name], dtype=np.float64).reshape(shape, order="F"))

    def I(name, *shape):
        return np.asfortranarray(np.asarray(a[name], dtype=np.int64).reshape(shape, order="F"))

    shapvt = F("shapvt", nge, npe, nd + 1)
    shapft = F("shapft", ngf, npf, nd)
    m = Master(nd=nd, porder=h["porder"], pgauss=2 * h["porder"], plocvl=F("plocvl", npe, nd),
               plocfc=F("plocfc", npf, nd - 1), perm=I("perm", npf, nfe) - 1, permgeom=I("permgeom", nmf, nfe) - 1,
               gpvl=F("gpvl", nge, nd), gwvl=F("gwvl", nge).ravel(), gpfc=F("gpfc", ngf, nd - 1),
               gwfc=F("gwfc", ngf).ravel(), shapvl=np.asfortranarray(np.transpose(shapvt, (1, 0, 2))),
               shapfc=np.asfortranarray(np.transpose(shapft, (1, 0, 2))), shapmv=F("shapmv", nge, nme, nd + 1),
               shapmf=F("shapmf", ngf, nmf, nd))
    m.shapvt, m.shapvg = shapvt, F("shapvg", npe, nge, nd + 1)
    m.shapvgdotshapvl = F("shapvgdotshapvl", npe * npe, nge, nd + 1)
    m.shapft, m.shapfg = shapft, F("shapfg", npf, ngf, nd)
    m.shapfgdotshapfc = F("shapfgdotshapfc", npf * npf, ngf, nd)

    nb, nch = h["nb"], h["nch"]
    flag = np.asarray(a["flag"])
    factor = np.asarray(a["factor"], dtype=np.float64)
    if physics is None:
        if "problem" not in a or len(a["problem"]) < 2:
            raise ValueError("physics is not stored in this layout: pass physics=")
        physics = int(a["problem"][1])
    ncd = h["ncd"]
    if av_col is None:
        av_col = nd if ncd > nd else -1
    app = App(physics=physics, nd=nd, ncu=h["ncu"], nc=h["nc"], param=[float(v) for v in a["param"]],
              bcm=[int(v) for v in a["bcm"]], bcs=np.asarray(a["bcs"], dtype=np.float64).reshape(nch, nb, order="F").T,
              localsolve=int(flag[5]) if flag.size > 5 else int(h["nc"] > h["ncu"]),
              tdep=int(flag[0]) if flag.size else 0,
              fc_u=float(factor[0]) if factor.size else 0.0, fc_q=float(factor[1]) if factor.size > 1 else 1.0,
              av_col=av_col)

    t2f = I("t2f", ne, nfe) - 1
    bf = I("bf", nfe, ne)
    if "f" in a:
        nvf = h["nvf"]
        f = I("f", h["nf"], nvf + 2).copy()
        f[:, :nvf] -= 1
        f[:, nvf] -= 1
        e2 = f[:, nvf + 1]
        f[:, nvf + 1] = np.where(e2 > 0, e2 - 1, np.where(e2 == 0, -1, e2))   # 0: boundary without an id
    else:
        ef = faces_from_t2f(t2f, bf)
        f = np.concatenate([-np.ones((ef.shape[0], nd), dtype=np.int64), ef], axis=1)
    mesh = Mesh(nd=nd, porder=h["morder"], p=np.zeros((0, nd)), t=-np.ones((ne, nd + 1), dtype=np.int64), f=f,
                t2f=t2f, bf=np.asfortranarray(bf.astype(np.int32)), dgnodes=F("dgnodes", nme, ncd, ne),
                elcon=I("elcon", npf * nfe, ne) - 1, nsiz=h["ndh"])
    UDG = F("UDG", npe, h["nc"], ne)
    UH = F("UH", nch, h["ndh"])
    return m, app, mesh, UDG, UH

"""A request's mesh connectivity built on the device, as the serving loop does
on admission (the reference builds it on the host in mesh/mkmesh/mkmesh.m):

  * faces and element-to-face map: hdgx_mesh_faces (mesh/mkmesh/mkt2f.m);
  * boundary tags of the unit square, setbndnbrs with the bndexpr of
    mesh/mkmesh_square.m (1: y = 0, 2: x = 1, 3: y = 1, 4: x = 0), and
    bf = f(t2f, end) per element face (kernel/hdg_assemble.m's bf);
  * trace-node numbering: hdgx_mesh_elcon (mesh/mkmesh/elconnectivities.m);
  * face-to-face table and face connectivity: hdgx_mesh_f2f (mkf2f.m);
  * the initial trace UH, the average of the adjacent elements' traces
    (util/inituhat.m:17-27), by an index-add on the device (each trace node
    has one or two contributions, so the sum is the same in any order).

Only the greedy face colouring of the block ILU(0) order runs on the host
(hdgx_color_faces, sequential by nature) on the downloaded face table.
Integer arrays equal hdg_amd/mesh.py's and hdg_amd/trace.py's host
construction; UH equals hdg_amd/problems.inituhat bit for bit."""
from dataclasses import dataclass

import numpy as np

from .runtime import Dev, HdgxError, _torch
from .trace import TraceSystem


@dataclass
class DeviceMesh:
    ne: int
    nf: int
    nsiz: int
    dgnodes: Dev       # (npv, ncd, ne)
    bf: Dev            # (nfe, ne) int32
    elcon: Dev         # (npf*nfe, ne) int32
    fconn: Dev         # (4, nf) int32
    f2f: Dev           # (2nfe-1, nf) int32
    f2f_host: np.ndarray


def square_boundary_tags(P, f, nf, nd=2):
    """-(boundary id) of every boundary face of the unit square and the
    second element elsewhere (setbndnbrs.m with mkmesh_square.m's bndexpr:
    the first of y = 0, x = 1, y = 1, x = 0 that holds for all the face's
    vertices).  P (nv, 2) device tensor, f the Dev of hdgx_mesh_faces."""
    torch = _torch()
    F = f.t.view(nf, nd + 2).long()
    e2 = F[:, nd + 1]
    xy = P[F[:, :nd]]                                # (nf, nd, 2)
    x, y = xy[..., 0], xy[..., 1]
    tol = 1e-6
    c = [(y < tol).all(1), (x > 1 - tol).all(1), (y > 1 - tol).all(1), (x < tol).all(1)]
    tag = torch.zeros(nf, dtype=torch.long, device=P.device)
    for k in (3, 2, 1, 0):                           # the first true test wins
        tag = torch.where(c[k], torch.full_like(tag, k + 1), tag)
    bnd = e2 < 0
    if bool((bnd & (tag == 0)).any()):
        raise HdgxError("a boundary face on no side of the unit square")
    return torch.where(bnd, -tag, e2)


def build(ctx, vertices, t, master, dgnodes, udg, ncu):
    """Connectivity, bf and the initial trace of a square-mesh request on the
    device.  vertices (nv, 2), t (ne, 3) host arrays; dgnodes (npv, ncd, ne)
    and udg (npv, nc, ne) host or Dev.  ctx's stream carries the work (the
    mesh entry points synchronise it).  Returns (DeviceMesh, UH Dev (ncu, nsiz),
    TraceSystem)."""
    torch = _torch()
    t = np.asarray(t)
    ne, nv = t.shape[0], np.asarray(vertices).shape[0]
    nd = t.shape[1] - 1
    nfe = nd + 1
    f, t2f, nf = ctx.mesh_faces(t, nvert=nv)
    dg = dgnodes if isinstance(dgnodes, Dev) else Dev.from_numpy(dgnodes)
    with torch.cuda.stream(ctx.stream):
        P = torch.from_numpy(np.ascontiguousarray(vertices, dtype=np.float64)).to("cuda", non_blocking=False)
        last = square_boundary_tags(P, f, nf, nd)
        T2F = t2f.t.view(ne, nfe).long()
        bfv = last[T2F]                                  # (ne, nfe)
        bf = Dev(torch.where(bfv < 0, bfv, bfv + 1).to(torch.int32).flatten(), (nfe, ne))
    elcon = ctx.mesh_elcon(dg, master.perm, f, t2f, nf)
    fconn, f2f = ctx.mesh_f2f(f, t2f, nf)
    npf = master.perm.shape[0]
    nsiz = nf * npf
    f2f_host = f2f.numpy()
    mesh = DeviceMesh(ne, nf, nsiz, dg, bf, elcon, fconn, f2f, f2f_host)
    UH = init_trace(ctx, master, elcon, udg, ncu, nsiz)
    ts = TraceSystem.from_device(ctx, ncu, npf, nfe, ne, nf, fconn, f2f, f2f_host, elcon)
    ctx.stream.synchronize()          # everything ready for other streams
    return mesh, UH, ts


def init_trace(ctx, master, elcon, udg, ncu, nsiz):
    """UH (ncu, nsiz): every trace node the average of its adjacent
    elements' traces (util/inituhat.m:17-27), on the device."""
    torch = _torch()
    u = udg if isinstance(udg, Dev) else Dev.from_numpy(udg)
    npv, nc, ne = u.shape
    perm = np.asarray(master.perm)
    nloc = perm.size
    with torch.cuda.stream(ctx.stream):
        U = u.t.view(ne, nc, npv)
        pidx = torch.from_numpy(perm.ravel(order="F").astype(np.int64)).to("cuda")
        vals = U[:, :ncu, :].index_select(2, pidx)                  # (ne, ncu, nloc)
        EL = elcon.t.view(ne, nloc).long().flatten()
        acc = torch.zeros((nsiz, ncu), dtype=torch.float64, device="cuda")
        acc.index_add_(0, EL, vals.permute(0, 2, 1).reshape(-1, ncu))
        cnt = torch.bincount(EL, minlength=nsiz).to(torch.float64)
        UH = (acc / cnt[:, None]).flatten()
    return Dev(UH, (ncu, nsiz))

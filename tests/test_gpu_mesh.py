"""Mesh connectivity on the device (hdgx_mesh_faces, hdgx_mesh_elcon,
hdgx_mesh_f2f) equals the host construction of hdg_amd/mesh.py and
hdg_amd/trace.py (mesh/mkmesh/mkt2f.m, elconnectivities.m:147-212, mkf2f.m),
and the CBSR pattern of hdgx_mesh_block_crs equals the oracle's elcon2entcon
restatement (block_crs.m), integer for integer: square and cube meshes,
jittered and re-oriented, one element, no element, and the error paths
(a face of three elements, a vertex id out of range, trace nodes that do not
meet)."""
from types import SimpleNamespace

import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.master import mkmaster  # noqa: E402
from hdg_amd.mesh import _orient, createnodes, cubemesh, hdgelcon, jitter, mkt2f, squaremesh  # noqa: E402
from hdg_amd.runtime import HdgxError  # noqa: E402
from hdg_amd.trace import face_connectivity, mkf2f  # noqa: E402
from hdg_amd.runtime import Dev  # noqa: E402
from tests.test_mkt2f_bruteforce import MESHES, mkt2f_loop  # noqa: E402
from oracle import trace_oracle as TO  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


def _check(ctx, p, t, master):
    nd = p.shape[1]
    f_ref, t2f_ref = mkt2f(t, nd)
    f, t2f, nf = ctx.mesh_faces(t)
    assert nf == f_ref.shape[0]
    assert np.array_equal(f.numpy().T, f_ref)
    assert np.array_equal(t2f.numpy().T, t2f_ref)
    dg = createnodes(p, t, master.plocvl)
    el_ref, nsiz = hdgelcon(dg, master.perm, f_ref, t2f_ref)
    el = ctx.mesh_elcon(dg, master.perm, f, t2f, nf)
    assert nsiz == nf * master.perm.shape[0]
    assert np.array_equal(el.numpy(), el_ref)
    ms = SimpleNamespace(f=f_ref, t2f=t2f_ref)
    fconn, f2f = ctx.mesh_f2f(f, t2f, nf)
    assert np.array_equal(fconn.numpy(), face_connectivity(ms))
    assert np.array_equal(f2f.numpy(), mkf2f(ms))
    rowpts, colind = ctx.mesh_block_crs(f2f)
    r_ref, c_ref = TO.entcon(t2f_ref)
    assert np.array_equal(rowpts.numpy(), r_ref)
    assert np.array_equal(colind.numpy(), c_ref)


@pytest.mark.parametrize("nx,ny,porder,amp", [(7, 5, 3, 0.2), (64, 64, 2, 0.0), (200, 150, 4, 0.1)])
def test_square(ctx, nx, ny, porder, amp):
    print(f"[gpu] device mesh: square {nx}x{ny} p{porder}", flush=True)
    p, t = squaremesh(nx + 1, ny + 1)
    if amp:
        p = jitter(p, 1.0 / max(nx, ny), amp, seed=nx)
    _check(ctx, p, t, mkmaster(2, porder))


@pytest.mark.parametrize("n,porder", [(3, 2), (12, 1), (16, 2)])
def test_cube(ctx, n, porder):
    print(f"[gpu] device mesh: cube {n} p{porder}", flush=True)
    p, t = cubemesh(n + 1, n + 1, n + 1)
    p = jitter(p, 1.0 / n, 0.2, seed=n)
    t = _orient(p, t)
    _check(ctx, p, t, mkmaster(3, porder))


def test_one_and_no_element(ctx):
    print("[gpu] device mesh: one triangle, one tetrahedron, empty", flush=True)
    _check(ctx, np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]), np.array([[0, 1, 2]]), mkmaster(2, 2))
    _check(ctx, np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]), np.array([[0, 1, 2, 3]]), mkmaster(3, 2))
    f, t2f, nf = ctx.mesh_faces(np.zeros((0, 3), dtype=np.int64), nvert=0)
    assert nf == 0


@pytest.mark.parametrize("name", list(MESHES))
def test_faces_match_reference_loop(ctx, name):
    """hdgx_mesh_faces against the literal restatement of mkt2f.m:58-80's
    first-visit loop (not against hdg_amd/mesh.py), integer for integer."""
    t = MESHES[name]()
    print(f"[gpu] device mesh vs mkt2f loop: {name} ne={t.shape[0]}", flush=True)
    f_ref, t2f_ref = mkt2f_loop(t, t.shape[1] - 1)
    f, t2f, nf = ctx.mesh_faces(t)
    assert nf == f_ref.shape[0]
    assert np.array_equal(f.numpy().T, f_ref)
    assert np.array_equal(t2f.numpy().T, t2f_ref)


def test_malformed_face_table(ctx):
    """An f whose element ids are 1-based (the reference's own numbering) or
    otherwise past ne is refused before any kernel indexes with it."""
    print("[gpu] device mesh: malformed f", flush=True)
    m = mkmaster(2, 2)
    p, t = squaremesh(4, 4)
    f, t2f, nf = ctx.mesh_faces(t)
    dg = createnodes(p, t, m.plocvl)
    bad = f.numpy().copy()
    bad[2:, :] += 1                              # element ids 1-based: the last element becomes ne
    fb = Dev.from_numpy(bad, np.int32)
    with pytest.raises(HdgxError, match="outside 0..ne-1"):
        ctx.mesh_elcon(dg, m.perm, fb, t2f, nf)
    with pytest.raises(HdgxError, match="outside 0..ne-1"):
        ctx.mesh_f2f(fb, t2f, nf)
    # a perm entry past npv
    perm = m.perm.copy()
    perm[0, 0] = m.npv
    with pytest.raises(HdgxError, match="perm entry"):
        ctx.mesh_elcon(dg, perm, f, t2f, nf)
    # the good tables still work afterwards
    el = ctx.mesh_elcon(dg, m.perm, f, t2f, nf)
    el_ref, _ = hdgelcon(dg, m.perm, *mkt2f(t, 2))
    assert np.array_equal(el.numpy(), el_ref)


def test_errors(ctx):
    print("[gpu] device mesh: error paths", flush=True)
    # three triangles on the edge (0, 1)
    with pytest.raises(HdgxError, match="more than two"):
        ctx.mesh_faces(np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]]))
    with pytest.raises(HdgxError, match="vertex id"):
        ctx.mesh_faces(np.array([[0, 1, 5]]), nvert=3)
    # the second element's copy of the shared face nodes moved off the first's
    m = mkmaster(2, 2)
    p, t = squaremesh(3, 3)
    f, t2f, nf = ctx.mesh_faces(t)
    dg = createnodes(p, t, m.plocvl)
    dg[:, :, 1] += 1e-6
    with pytest.raises(HdgxError, match="do not match"):
        ctx.mesh_elcon(dg, m.perm, f, t2f, nf)


def test_block_crs_errors(ctx):
    """A face table whose slot 0 is not the face, or with an entry past nf,
    raises before colind is written; no faces gives rowpts = [0]."""
    print("[gpu] device block_crs error paths", flush=True)
    p, t = squaremesh(4, 4)
    f, t2f, nf = ctx.mesh_faces(t)
    _, f2f = ctx.mesh_f2f(f, t2f, nf)
    bad = f2f.numpy().copy()
    bad[2, 5] = nf
    with pytest.raises(HdgxError, match="block_crs"):
        ctx.mesh_block_crs(Dev.from_numpy(bad, "int32"))
    bad = f2f.numpy().copy()
    bad[0, 3] = 4
    with pytest.raises(HdgxError, match="block_crs"):
        ctx.mesh_block_crs(Dev.from_numpy(bad, "int32"))
    rowpts, colind = ctx.mesh_block_crs(Dev.empty((5, 0), "int32"))
    assert rowpts.numpy().tolist() == [0] and colind.shape == (0,)

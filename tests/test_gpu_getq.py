"""GPU parity of hdgx_getq (kernel/getq.m) against the getq oracle, through the
C ABI: QDG and qq to 1e-12 normwise relative, the in-place form over UDG's q
columns equal to the separate one bit for bit, the closed form q = -grad u /
fc_q for a polynomial u, and the argument checks."""
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.gpu_util import relerr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.master import mkshape  # noqa: E402
from hdg_amd.problems import euler2d_square, euler3d_cube, ns2d_square  # noqa: E402
from hdg_amd.runtime import Dev, HdgxError  # noqa: E402
from oracle import getq_oracle as G  # noqa: E402
from tests.test_getq_oracle import own_trace, poly_field  # noqa: E402

CASES = {
    "tri_p1": lambda: euler2d_square(3, 1, seed=41),
    "tri_p2": lambda: ns2d_square(3, 2, seed=42),
    "tri_p3": lambda: ns2d_square(4, 3, seed=43),
    "tri_p4": lambda: euler2d_square(3, 4, seed=44),
    "tet_p1": lambda: euler3d_cube(2, 1, seed=45),
    "tet_p2": lambda: euler3d_cube(2, 2, seed=46),
}


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


def _state(P, seed):
    """UDG with nc = ncu*(nd+1) (random q columns, which getq must not read),
    a perturbed gathered trace and a random SH."""
    rng = np.random.default_rng(seed)
    nd, ncu = P.master.nd, P.app.ncu
    u = P.udg[:, :ncu, :]
    q = rng.standard_normal((u.shape[0], ncu * nd, u.shape[2]))
    udg = np.asfortranarray(np.concatenate([u, q], axis=1))
    uh = np.asfortranarray(P.uh * (1.0 + 1e-2 * rng.standard_normal(P.uh.shape)))
    sh = np.asfortranarray(rng.standard_normal(udg.shape))
    return udg, uh, sh


@pytest.mark.parametrize("case", list(CASES))
def test_getq_parity(ctx, case):
    P = CASES[case]()
    m = P.master
    print(f"[gpu] getq {case} ne={P.ne}", flush=True)
    udg, uh, sh = _state(P, 7)
    fc_q = 1.7
    shapnv = mkshape(m.porder, m.plocvl, m.plocvl)
    ref_q, ref_qq = G.getq(m, P.mesh.dgnodes, udg, uh, SH=sh, fc_q=fc_q, shapnv=shapnv)
    q, qq = ctx.getq(m, P.mesh.dgnodes, udg, uh, sh=sh, fc_q=fc_q, shapnv=shapnv)
    q0 = ctx.getq(m, P.mesh.dgnodes, udg, uh, fc_q=fc_q)
    ctx.stream.synchronize()
    assert relerr(q.numpy(), ref_q) < 1e-12
    assert relerr(qq.numpy(), ref_qq) < 1e-12
    assert relerr(q0.numpy(), G.getq(m, P.mesh.dgnodes, udg, uh, fc_q=fc_q)) < 1e-12

    # in place over UDG(:, ncu+1:end, :): u columns untouched, q columns equal
    # to the separate output bit for bit
    U = Dev.from_numpy(udg)
    ctx.getq(m, P.mesh.dgnodes, U, uh, sh=sh, fc_q=fc_q, inplace=True)
    ctx.stream.synchronize()
    got = U.numpy()
    ncu = P.app.ncu
    assert np.array_equal(got[:, :ncu, :], udg[:, :ncu, :])
    assert np.array_equal(got[:, ncu:, :], q.numpy())


@pytest.mark.parametrize("case", ["tri_p3", "tet_p2"])
def test_getq_closed_form(ctx, case):
    P = CASES[case]()
    m, nd, ncu = P.master, P.master.nd, P.app.ncu
    u, gu = poly_field(P.mesh.dgnodes, nd, m.porder, ncu, seed=5)
    udg = np.asfortranarray(np.concatenate([u, np.zeros((u.shape[0], ncu * nd, u.shape[2]))], axis=1))
    uhe = own_trace(u, np.asarray(m.perm))
    q, qq = ctx.getq(m, P.mesh.dgnodes, udg, uhe, fc_q=2.0, shapnv=mkshape(m.porder, m.plocvl, m.plocvl))
    ctx.stream.synchronize()
    want = -gu.reshape(u.shape[0], ncu * nd, u.shape[2], order="F")
    scale = np.abs(want).max()
    assert np.abs(2.0 * q.numpy() - want).max() <= 1e-11 * scale
    assert np.abs(qq.numpy() - want).max() <= 1e-11 * scale


def test_getq_errors(ctx):
    P = CASES["tri_p2"]()
    m = P.master
    udg, uh, _ = _state(P, 1)
    with pytest.raises(HdgxError, match="fc_q"):
        ctx.getq(m, P.mesh.dgnodes, udg, uh, fc_q=0.0)
    with pytest.raises(HdgxError, match="inplace"):
        ctx.getq(m, P.mesh.dgnodes, udg, uh, inplace=True)
    out = ctx.getq(m, P.mesh.dgnodes, udg[:, :, :0].copy(order="F"), uh[:, :, :0].copy(order="F"))
    assert out.shape[2] == 0

"""HDGX_HOST_PTRS: the element-level entry points called with host (numpy)
arrays give, bit for bit, what they give on device arrays -- outputs, the
in-place q of getq, and with HDGX_INPLACE the inputs the MEX files overwrite."""
import ctypes

import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.gpu_util import perturb

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.app import MasterC  # noqa: E402
from hdg_amd.problems import euler2d_square, ns2d_square  # noqa: E402
from hdg_amd.runtime import INPLACE, Dev, lib  # noqa: E402

HOST_PTRS = 2


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


def hp(a):
    assert a.flags.f_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def zeros(*shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


def _host(ctx, fn):
    ctx.set_flags(HOST_PTRS)
    try:
        rc = fn()
    finally:
        ctx.set_flags(0)
    ctx._check(rc)


def test_elemental_and_matrices(ctx):
    P = perturb(ns2d_square(3, 3, seed=91, av_frac=0.3), 5)
    m, a = P.master, P.app
    print("[gpu] host pointers: elemental ns2d p3", flush=True)
    ref = [x.numpy() for x in ctx.elemental(m, a, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)]
    npv, ncu, nc, ne, ndf, nd = m.npv, a.ncu, a.nc, P.ne, m.ndf, m.nd
    nh = ncu * ndf
    out = [zeros(nh, nh, ne), zeros(nh, ne), zeros(npv, nc, ne), zeros(npv, nc, nh, ne), zeros(ne, dtype=np.int32)]
    mc, ca = MasterC(m), a.to_c()
    bf = np.asfortranarray(P.mesh.bf.astype(np.int32))
    ins = [np.asfortranarray(x) for x in (P.mesh.dgnodes, P.udg, P.uh, P.sh)]
    _host(ctx, lambda: lib().hdgx_elemental(ctx.h, ctypes.byref(mc.c), ctypes.byref(ca), ne, hp(ins[0]), hp(bf),
                                            hp(ins[1]), hp(ins[2]), hp(ins[3]), *[hp(o) for o in out]))
    for g, r in zip(out, ref):
        assert np.array_equal(g.reshape(r.shape, order="F"), r)
    mats = ctx.elemental_matrices(m, a, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    keys = ("BD", "F", "GK", "H", "Ru", "Rh", "M", "C", "E", "Rq")
    hm = {k: zeros(*mats[k].shape) for k in keys}
    _host(ctx, lambda: lib().hdgx_elemental_matrices(ctx.h, ctypes.byref(mc.c), ctypes.byref(ca), ne, hp(ins[0]),
                                                     hp(bf), hp(ins[1]), hp(ins[2]), hp(ins[3]),
                                                     *[hp(hm[k]) for k in keys]))
    for k in keys:
        assert np.array_equal(hm[k], mats[k].numpy()), k
    Ru, Rh, Rq = (zeros(npv, ncu, ne), zeros(ncu, ndf, ne), zeros(npv, ncu, nd, ne))
    _host(ctx, lambda: lib().hdgx_residual(ctx.h, ctypes.byref(mc.c), ctypes.byref(ca), ne, hp(ins[0]), hp(bf),
                                           hp(ins[1]), hp(ins[2]), hp(ins[3]), hp(Ru), hp(Rh), hp(Rq)))
    r = ctx.residual(m, a, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    for g, d in zip((Ru, Rh, Rq), r):
        assert np.array_equal(g, d.numpy())
    # getq in place over the host UDG's q columns
    U = ins[1].copy(order="F")
    qref = ctx.getq(m, P.mesh.dgnodes, P.udg, P.uh, fc_q=a.fc_q).numpy()
    qptr = ctypes.c_void_p(U.ctypes.data + 8 * npv * ncu)
    _host(ctx, lambda: lib().hdgx_getq(ctx.h, ctypes.byref(mc.c), ncu, nc, ins[0].shape[1], ne, hp(ins[0]), hp(U),
                                       hp(ins[2]), None, a.fc_q, qptr, npv * nc, None, None))
    assert np.array_equal(U[:, ncu:, :], qref)
    assert np.array_equal(U[:, :ncu, :], ins[1][:, :ncu, :])


def test_schur_inplace_on_host(ctx, oracle):
    P = perturb(ns2d_square(3, 2, seed=92, av_frac=0.3), 5)
    print("[gpu] host pointers: schur_uq with HDGX_INPLACE", flush=True)
    mats = oracle.elemental_matrices(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    keys = ("M", "C", "E", "BD", "F", "GK", "H", "Rq", "Ru", "Rh")
    dev = [Dev.from_numpy(mats[k]) for k in keys]
    ctx.set_flags(INPLACE)
    try:
        ref = [o.numpy() for o in ctx.schur_uq(*dev)]
    finally:
        ctx.set_flags(0)
    host = [np.asfortranarray(mats[k].copy()) for k in keys]
    npv, ne = P.master.npv, P.ne
    ndf, nd = P.master.ndf, P.master.nd
    ncu, nc = P.app.ncu, P.app.nc
    out = [zeros(npv, nc, ne), zeros(npv, nc, ncu * ndf, ne), zeros(ncu * ndf, ncu * ndf, ne), zeros(ncu * ndf, ne),
           zeros(ne, dtype=np.int32)]
    ctx.set_flags(INPLACE | HOST_PTRS)
    try:
        rc = lib().hdgx_schur_uq(ctx.h, nd, npv, ncu, nc, ndf, ne, *[hp(x) for x in host], *[hp(o) for o in out])
    finally:
        ctx.set_flags(0)
    ctx._check(rc)
    for g, r in zip(out, ref):
        assert np.array_equal(g.reshape(r.shape, order="F"), r)
    for k, h, d in zip(keys, host, dev):
        assert np.array_equal(h, d.numpy()), k


def test_schur_u_on_host(ctx, oracle):
    P = perturb(euler2d_square(3, 2, seed=93, alpha=0.3), 5)
    mats = oracle.elemental_matrices(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    keys = ("BD", "F", "GK", "H", "Ru", "Rh")
    ref = [o.numpy() for o in ctx.schur_u(*(mats[k] for k in keys))]
    host = [np.asfortranarray(mats[k]) for k in keys]
    npv, ne, ndf, ncu = P.master.npv, P.ne, P.master.ndf, P.app.ncu
    out = [zeros(npv, ncu, ne), zeros(npv, ncu, ncu * ndf, ne), zeros(ncu * ndf, ncu * ndf, ne), zeros(ncu * ndf, ne),
           zeros(ne, dtype=np.int32)]
    _host(ctx, lambda: lib().hdgx_schur_u(ctx.h, npv, ncu, ncu, ndf, ne, *[hp(x) for x in host],
                                          *[hp(o) for o in out]))
    for g, r in zip(out, ref):
        assert np.array_equal(g.reshape(r.shape, order="F"), r)

"""hdgx_elemental_noalloc, the elemental entry point that allocates nothing:
through the C ABI its outputs equal hdgx_elemental's bit for bit (the same
kernels), on two calls in a row -- a large NS p = 3 batch that runs in several
chunks of its fixed scratch array, then a smaller Euler p = 2 input with
another master -- and the device's free memory does not move across calls."""
import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.problems import euler2d_square, ns2d_square, with_bdf1  # noqa: E402

NAMES = ("AE", "FE", "DUDG", "DUDG_DUH", "status")


def _dev_case(P):
    from hdg_amd.runtime import Dev
    npv, ncu, nc = P.master.npv, P.app.ncu, P.app.nc
    ndf, ne = P.master.npf * P.master.nfe, P.ne
    ins = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
           Dev.from_numpy(P.uh), Dev.from_numpy(P.sh))

    def outs():
        return (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
                Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
    return ins, outs


def test_noalloc_equals_elemental_on_two_calls():
    import torch
    from hdg_amd.runtime import Context
    ctx = Context(0)
    # 2 x 36 x 36 NS p3 elements: 124 KB of element matrices each, so the 127 MiB
    # workspace part of the scratch array takes about 1000 at a time
    big = with_bdf1(ns2d_square(36, 3, seed=81, av_frac=0.2), 0.05, seed=8)
    small = euler2d_square(6, 2, seed=82, alpha=0.3)
    cases = []
    for P in (big, small):
        ins, outs = _dev_case(P)
        ref = ctx.elemental(P.master, P.app, *ins, out=outs())
        cases.append((P, ins, outs(), ref))
    torch.cuda.synchronize()
    # warm-up: the first call resolves nothing new (the scratch address was
    # taken at context creation) but loads the tables of the first master
    P, ins, out, ref = cases[0]
    ctx.elemental_noalloc(P.master, P.app, *ins, out=out)
    torch.cuda.synchronize()
    for P, ins, out, ref in cases:
        print(f"[gpu] elemental_noalloc {P.name} ne={P.ne}", flush=True)
        for o in out:
            o.t.zero_()
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info(0)[0]
        ctx.elemental_noalloc(P.master, P.app, *ins, out=out)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info(0)[0]
        assert free1 == free0, (free0, free1)
        for name, a, b in zip(NAMES, out, ref):
            assert np.array_equal(a.numpy(), b.numpy()), (P.name, name)
        assert np.all(out[4].numpy() == 0)


def test_noalloc_fp32_ir_and_refusals():
    import torch
    from hdg_amd.runtime import Context, HdgxError
    ctx = Context(0)
    P = ns2d_square(5, 3, seed=83, av_frac=0.3)
    ins, outs = _dev_case(P)
    print(f"[gpu] elemental_noalloc FP32+IR {P.name} ne={P.ne}", flush=True)
    ctx.set_precision(1)
    try:
        ref = ctx.elemental(P.master, P.app, *ins, out=outs())
        got = ctx.elemental_noalloc(P.master, P.app, *ins, out=outs())
        torch.cuda.synchronize()
    finally:
        ctx.set_precision(0)
    for name, a, b in zip(NAMES, got, ref):
        assert np.array_equal(a.numpy(), b.numpy()), name
    ctx.set_flags(2)                      # HDGX_HOST_PTRS: staging would allocate
    try:
        with pytest.raises(HdgxError, match="device arrays only"):
            ctx.elemental_noalloc(P.master, P.app, *ins, out=outs())
    finally:
        ctx.set_flags(0)

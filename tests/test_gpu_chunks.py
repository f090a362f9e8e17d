"""The element-matrix workspace chunking of hdgx_elemental and
hdgx_elemental_batch (the reference blocks its element loop the same way,
kernel/hdg_elemental.m:9-13,45): a workspace limit that forces three or more
chunks, the last one ragged, gives the same bits as one chunk.  In the batch
the chunk boundaries fall inside segments, so the per-chunk offsets into every
request's arrays are exercised."""
import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.problems import euler2d_square, euler3d_cube, ns2d_square, with_bdf1  # noqa: E402


def element_bytes(P):
    """Workspace bytes of one element: BD, F, GK, H, Ru, Rh and, with q, M, C,
    E, Rq (hdgx_elemental's per-element workspace)."""
    m, a = P.master, P.app
    npv, ncu, nc, nd = m.npv, a.ncu, a.nc, m.nd
    ndf = m.npf * m.nfe
    nu, nh, ncol = npv * ncu, ncu * ndf, npv * nc
    w = nu * ncol + nu * nh + nh * ncol + nh * nh + nu + nh
    if a.localsolve:
        w += npv * npv + npv * npv * nd + npv * ndf * nd + npv * ncu * nd
    return 8 * w


def chunks_of(ne, cap_elems):
    """The chunk sizes hdgx_elemental runs for a limit of cap_elems elements:
    as few chunks as fit, of nearly equal size (the last the smallest)."""
    nch = -(-ne // cap_elems)
    c = -(-ne // nch)
    return [min(c, ne - i) for i in range(0, ne, c)]


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    c = Context(0)
    yield c
    c.set_workspace_limit(0)


CASES = {
    "euler3d_p2": (lambda: euler3d_cube(3, 2, seed=41, alpha=0.3), 50),     # 162 elements: 41, 41, 41, 39
    "ns2d_p3": (lambda: with_bdf1(ns2d_square(5, 3, seed=42, av_frac=0.3), 0.05, seed=2), 16),  # 50: 13, 13, 13, 11
    "euler2d_p4": (lambda: euler2d_square(4, 4, seed=43, alpha=0.3), 9),   # 32: 8, 8, 8, 8
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", [0, 1])
def test_elemental_chunked_equals_one_chunk(ctx, case, prec):
    make, cap = CASES[case]
    P = make()
    sizes = chunks_of(P.ne, cap)
    print(f"[gpu] chunked elemental {case} prec={prec} ne={P.ne} chunks={sizes}", flush=True)
    assert len(sizes) >= 3
    args = (P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    ctx.set_precision(prec)
    try:
        ctx.set_workspace_limit(0)
        one = [o.numpy() for o in ctx.elemental(*args)]
        ctx.set_workspace_limit(cap * element_bytes(P) + element_bytes(P) // 2)
        many = [o.numpy() for o in ctx.elemental(*args)]
        ctx.stream.synchronize()
    finally:
        ctx.set_precision(0)
        ctx.set_workspace_limit(0)
    for name, a, b in zip(("AE", "FE", "DUDG", "DUDG_DUH", "status"), many, one):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("phys", ["ns", "euler3d"])
def test_batch_chunked_equals_one_chunk(ctx, phys):
    from hdg_amd.runtime import Dev
    if phys == "ns":
        probs = [with_bdf1(ns2d_square(n, 2, seed=60 + n, minf=m, av_frac=0.3), 0.05, seed=n)
                 for n, m in ((3, 0.25), (5, 0.5), (4, 0.65))]            # 18, 50, 32 elements
        cap = 30                                                          # chunks of 25: 25 | 50 | 75 fall inside segments
    else:
        probs = [euler3d_cube(n, 2, seed=70 + n, minf=m, alpha=0.3) for n, m in ((2, 0.3), (3, 0.5), (1, 0.45))]
        cap = 70                                                          # 48, 162, 6 elements: chunks of 54
    total = sum(P.ne for P in probs)
    print(f"[gpu] chunked batch {phys} ne={[P.ne for P in probs]} chunks={chunks_of(total, cap)}", flush=True)
    assert len(chunks_of(total, cap)) >= 3
    master = probs[0].master

    def run():
        segs, outs = [], []
        for P in probs:
            npv, ncu, nc = P.master.npv, P.app.ncu, P.app.nc
            ndf, ne = P.master.npf * P.master.nfe, P.ne
            dev = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
                   Dev.from_numpy(P.uh), Dev.from_numpy(P.sh))
            out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
                   Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
            segs.append((P.app, *dev, out))
            outs.append(out)
        ctx.elemental_batch(master, segs)
        ctx.stream.synchronize()
        return [[o.numpy() for o in out] for out in outs]

    try:
        ctx.set_workspace_limit(0)
        one = run()
        ctx.set_workspace_limit(cap * element_bytes(probs[0]) + element_bytes(probs[0]) // 2)
        many = run()
    finally:
        ctx.set_workspace_limit(0)
    for P, a_seg, b_seg in zip(probs, many, one):
        for name, a, b in zip(("AE", "FE", "DUDG", "DUDG_DUH", "status"), a_seg, b_seg):
            assert np.array_equal(a, b), (P.name, name)
    # and each request's outputs are those of its own call: bit for bit in 2D;
    # for tetrahedra the batched quadrature kernel is a separate compilation
    # whose FMA contraction differs from the single call's in a few last bits
    # of F (DUDG and FE stay equal), hence a rounding bar there
    from tests.gpu_util import relerr
    for P, got in zip(probs, one):
        ref = ctx.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
        for name, a, b in zip(("AE", "FE", "DUDG", "DUDG_DUH", "status"), got, ref):
            if P.master.nd == 2 or name == "status":
                assert np.array_equal(a, b.numpy()), (P.name, name)
            else:
                assert relerr(a, b.numpy()) < 1e-13, (P.name, name)


def test_kernel_timing_counts_each_chunk(ctx):
    """hdgx_set_timing brackets every chunk's quadrature and elimination with
    events: one launch pair per chunk, positive times, outputs unchanged."""
    P = euler3d_cube(3, 2, seed=41, alpha=0.3)
    cap = 50
    sizes = chunks_of(P.ne, cap)
    print(f"[gpu] kernel timing euler3d_p2 ne={P.ne} chunks={sizes}", flush=True)
    args = (P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    try:
        ctx.set_workspace_limit(cap * element_bytes(P) + element_bytes(P) // 2)
        plain = [o.numpy() for o in ctx.elemental(*args)]
        ctx.set_timing(True)
        timed = [o.numpy() for o in ctx.elemental(*args)]
        tq, ts, n = ctx.kernel_times()
        ctx.set_timing(False)
        assert ctx.kernel_times() == (0.0, 0.0, 0)
    finally:
        ctx.set_timing(False)
        ctx.set_workspace_limit(0)
    assert n == len(sizes)
    assert tq > 0 and ts > 0
    for a, b in zip(timed, plain):
        assert np.array_equal(a, b)

"""GPU tests of the batched elemental launch and the serving loop: a batch
reproduces each request's own elemental call, a served request's result is
the same alone and in a mix of requests, and one served BDF1 Newton step
matches the oracle's."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests.conftest import has_gpu
from tests.gpu_util import relerr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.problems import euler2d_square, gather_uh, inituhat, ns2d_square, with_bdf1  # noqa: E402
from oracle import trace_oracle as TO  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


@pytest.mark.parametrize("phys", ["ns", "euler"])
def test_elemental_batch_matches_each_request_alone(ctx, phys):
    from hdg_amd.runtime import Dev
    print(f"[gpu] elemental batch {phys}", flush=True)
    if phys == "ns":
        probs = [with_bdf1(ns2d_square(n, 3, seed=50 + n, minf=m, av_frac=0.3), 0.05, seed=n)
                 for n, m in ((2, 0.25), (3, 0.5), (4, 0.65))]
    else:
        probs = [euler2d_square(n, 2, seed=60 + n, minf=m, alpha=0.3) for n, m in ((2, 0.3), (5, 0.6), (3, 0.45))]
    master = probs[0].master
    segs, outs, singles = [], [], []
    for P in probs:
        npv, ncu, nc = P.master.npv, P.app.ncu, P.app.nc
        ndf = P.master.npf * P.master.nfe
        ne = P.ne
        dev = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
               Dev.from_numpy(P.uh), Dev.from_numpy(P.sh))
        out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
               Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
        segs.append((P.app, *dev, out))
        outs.append(out)
        singles.append(ctx.elemental(master, P.app, *dev))
    ctx.elemental_batch(master, segs)
    ctx.stream.synchronize()
    for out, ref in zip(outs, singles):
        for a, b in zip(out, ref):
            assert np.array_equal(a.numpy(), b.numpy())


def _small_workload(nreq=8, seed=7):
    from hdg_amd.serve import RequestSpec
    rng = np.random.default_rng(seed)
    return [RequestSpec(i, 0.0, int(rng.integers(3, 7)), int(1 + i % 4), float(rng.uniform(0.2, 0.7)),
                        int(rng.integers(0, 1 << 30)), 0.05, "ns", iters=1 + (i % 3 == 0)) for i in range(nreq)]


def test_served_result_is_the_same_alone_and_in_the_mix():
    from hdg_amd.serve import PayloadFactory, Server
    specs = _small_workload()
    fac = PayloadFactory()
    print("[gpu] serve mix", flush=True)
    # a batching window far longer than the admissions: the 8 requests, of
    # orders 1..4, are admitted together and their first elemental call is
    # shared across orders
    srv = Server(workers=3, slots=16, batch_window=0.5)
    stats, done = srv.run(specs, factory=fac, realtime=False, keep_state=True)
    assert stats["requests"] == len(specs)
    mix = {s.spec.rid: (s.state.UDG.numpy(), s.state.UH.numpy()) for s in done}
    print(f"[gpu]   {stats['elemental_batches']} batches, {stats['mean_requests_per_batch']:.2f} requests each, "
          f"{stats['mixed_order_batches']} over several orders", flush=True)
    assert stats["elemental_batches"] < sum(s.iters for s in specs)
    assert stats["mixed_order_batches"] >= 1
    for rid in (0, 3, 5):
        print(f"[gpu] serve alone {rid}", flush=True)
        _, one = Server(workers=1, slots=1).run([specs[rid]], factory=fac, realtime=False, keep_state=True)
        u, h = one[0].state.UDG.numpy(), one[0].state.UH.numpy()
        assert relerr(mix[rid][0], u) <= 1e-12 and relerr(mix[rid][1], h) <= 1e-12


def test_tight_budget_admits_one_at_a_time_with_the_same_results():
    from hdg_amd.serve import PayloadFactory, Server
    specs = _small_workload(5, seed=9)
    fac = PayloadFactory()
    print("[gpu] serve budget", flush=True)
    _, a = Server(workers=2, slots=8).run(specs, factory=fac, realtime=False, keep_state=True)
    _, b = Server(workers=2, slots=8, budget_bytes=1).run(specs, factory=fac, realtime=False, keep_state=True)
    ra = {s.spec.rid: s.state.UH.numpy() for s in a}
    rb = {s.spec.rid: s.state.UH.numpy() for s in b}
    for k in ra:
        assert relerr(ra[k], rb[k]) <= 1e-12


@pytest.mark.parametrize("p", [2, 3])
def test_served_bdf1_step_matches_the_oracle(oracle, p):
    """One served Newton step of a BDF1 request vs the oracle: elemental on
    the CPU, direct sparse trace solve, local update (hdg_solve.m:112-146)."""
    from hdg_amd.serve import PayloadFactory, RequestSpec, Server
    spec = RequestSpec(0, 0.0, 4, p, 0.45, 1234, 0.05, "ns", iters=1)
    fac = PayloadFactory()
    pl = fac.build(spec)
    print(f"[gpu] served step vs oracle p={p}", flush=True)
    _, done = Server(workers=1, slots=1, gmres_tol=1e-13).run([spec], factory=fac, realtime=False, keep_state=True)
    u_gpu, uh_gpu = done[0].state.UDG.numpy(), done[0].state.UH.numpy()
    m, app = pl["master"], pl["app"]
    mesh = fac.host_mesh(spec.n, spec.p)
    udg = pl["udg"].copy()
    UH = inituhat(m, mesh.elcon, udg, app.ncu, mesh.nsiz)
    uh = gather_uh(UH, mesh.elcon)
    AE, FE, DUDG, DUDG_DUH, st = oracle.elemental(m, app, mesh.dgnodes, mesh.bf, udg, uh, pl["sh"])
    A, b = TO.sparse_system(AE, FE, mesh.elcon, app.ncu, mesh.nsiz)
    DUH = spla.spsolve(A.tocsc(), b).reshape(app.ncu, -1, order="F")
    udg[:, :, :] += TO.local_update(DUDG, DUDG_DUH, DUH, mesh.elcon)
    UH = UH + DUH
    e_uh, e_u = relerr(uh_gpu.reshape(UH.shape, order="F"), UH), relerr(u_gpu, udg)
    print(f"[gpu]   UH {e_uh:.2e} UDG {e_u:.2e}", flush=True)
    assert e_uh < 1e-12 and e_u < 1e-12


def test_realtime_serving_reports_latency():
    from hdg_amd.serve import Server, make_workload
    specs = make_workload(12, seed=3, rate=200.0, ne_lo=200, ne_hi=3000)
    print("[gpu] realtime serve", flush=True)
    stats, done = Server(workers=2, slots=8).run(specs)
    assert stats["requests"] == 12
    assert 0 < stats["p50_ms"] <= stats["p99_ms"]


@pytest.mark.parametrize("n,p", [(3, 1), (6, 2), (9, 3), (5, 4), (40, 3)])
def test_device_connectivity_equals_host_construction(n, p):
    """The serving loop's on-admission connectivity (hdg_amd.devmesh, on the
    device) equals the host construction (hdg_amd/mesh.py, trace.py,
    problems.inituhat): bf, elcon, fconn, f2f and the face colouring integer
    for integer, the initial trace bit for bit."""
    from hdg_amd import devmesh
    from hdg_amd.runtime import Context
    from hdg_amd.serve import PayloadFactory, RequestSpec
    from hdg_amd.trace import color_faces, face_connectivity, mkf2f
    fac = PayloadFactory()
    pl = fac.build(RequestSpec(0, 0.0, n, p, 0.4, 77, 0.05, "ns"))
    m, app = pl["master"], pl["app"]
    print(f"[gpu] device connectivity n={n} p={p}", flush=True)
    ctx = Context(0)
    dm, UH, ts = devmesh.build(ctx, pl["vertices"], pl["t"], m, pl["dgnodes"], pl["udg"], app.ncu)
    mesh = fac.host_mesh(n, p)
    assert np.array_equal(dm.bf.numpy(), mesh.bf)
    assert np.array_equal(dm.elcon.numpy(), mesh.elcon)
    assert np.array_equal(dm.fconn.numpy(), face_connectivity(mesh))
    f2f = mkf2f(mesh)
    assert np.array_equal(dm.f2f.numpy(), f2f)
    assert np.array_equal(ts.face_color, color_faces(f2f)[3])
    assert np.array_equal(UH.numpy(), inituhat(m, mesh.elcon, pl["udg"], app.ncu, mesh.nsiz))


def test_mixed_order_batch_equals_each_order_batch():
    """hdgx_elemental_batch_mixed over NS requests of orders 1..4 (interleaved)
    gives each request the bits of its own order's hdgx_elemental_batch."""
    from hdg_amd.runtime import Dev
    print("[gpu] mixed-order batch", flush=True)
    probs = [with_bdf1(ns2d_square(n, p, seed=90 + p, minf=m, av_frac=0.3), 0.05, seed=p)
             for n, p, m in ((3, 2, 0.3), (2, 4, 0.5), (4, 1, 0.6), (3, 3, 0.4), (2, 2, 0.25), (5, 1, 0.45))]

    def arrays(P):
        npv, ncu, nc, ndf, ne = P.master.npv, P.app.ncu, P.app.nc, P.master.npf * P.master.nfe, P.ne
        dev = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
               Dev.from_numpy(P.uh), Dev.from_numpy(P.sh))
        out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
               Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
        return (P.app, *dev, out)

    from hdg_amd.runtime import Context
    ctx = Context(0)
    segs = [arrays(P) for P in probs]
    ctx.elemental_batch_mixed([P.master for P in probs], segs)
    for P, seg in zip(probs, segs):
        one = arrays(P)
        ctx.elemental_batch(P.master, [one])
        for a, b in zip(seg[6], one[6]):
            assert np.array_equal(a.numpy(), b.numpy()), P.name

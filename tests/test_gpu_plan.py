"""GPU tests of the mixed-order launch plans (hdgx_plan_*): every mode, launched
directly and replayed from its captured hipGraph several times with eager work
in between, reproduces each segment's own hdgx_elemental result."""
import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.problems import euler2d_square, euler3d_cube, ns2d_square, with_bdf1  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


def _segments(probs):
    from hdg_amd.runtime import Dev
    masters, segs, singles = [], [], []
    for P in probs:
        npv, ncu, nc = P.master.npv, P.app.ncu, P.app.nc
        ndf = P.master.npf * P.master.nfe
        ne = P.ne
        dev = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
               Dev.from_numpy(P.uh), Dev.from_numpy(P.sh))
        out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
               Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
        masters.append(P.master)
        segs.append((P.app, *dev, out))
    return masters, segs


FAMILIES = {
    "euler2d": lambda: [euler2d_square(n, p, seed=70 + p, minf=m, alpha=0.3)
                        for n, p, m in ((4, 1, 0.3), (3, 3, 0.5), (5, 2, 0.6), (3, 4, 0.45), (2, 1, 0.55))],
    "ns2d": lambda: [with_bdf1(ns2d_square(n, p, seed=80 + p, minf=m, av_frac=0.3), 0.05, seed=p)
                     for n, p, m in ((3, 2, 0.3), (2, 4, 0.5), (3, 1, 0.4), (2, 3, 0.6))],
    "euler3d": lambda: [euler3d_cube(1, p, seed=90 + p, alpha=0.3) for p in (1, 2, 1)],
}


@pytest.mark.parametrize("family", list(FAMILIES))
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_plan_matches_each_segment_alone(ctx, family, mode):
    probs = FAMILIES[family]()
    print(f"[gpu] plan {family} mode {mode}", flush=True)
    masters, segs = _segments(probs)
    ref = [ctx.elemental(P.master, P.app, *seg[1:6]) for P, seg in zip(probs, segs)]
    ctx.stream.synchronize()
    refs = [[a.numpy() for a in r] for r in ref]
    plan = ctx.plan(masters, segs, mode, 64 if mode == 1 else 0)

    def check():
        ctx.stream.synchronize()
        for seg, r in zip(segs, refs):
            for got, want in zip(seg[6], r):
                g = got.numpy()
                assert np.array_equal(g, want) or np.abs(g - want).max() <= 1e-14 * max(1.0, np.abs(want).max())

    plan.launch()
    check()
    plan.capture()
    for rep in range(3):
        for seg in segs:                       # eager work between the replays
            for o in seg[6]:
                o.t.zero_()
        ctx.stream.synchronize()
        plan.launch()
        check()
    plan.close()


def test_autotuner_picks_a_plan_that_is_correct(ctx):
    from hdg_amd.autotune import tune
    probs = FAMILIES["euler2d"]()
    masters, segs = _segments(probs)
    ref = [ctx.elemental(P.master, P.app, *seg[1:6]) for P, seg in zip(probs, segs)]
    ctx.stream.synchronize()
    refs = [[a.numpy() for a in r] for r in ref]
    plan, timings = tune(ctx, masters, segs, reps=2)
    assert len(timings) >= 3 and all(t > 0 for t in timings.values())
    for seg in segs:
        for o in seg[6]:
            o.t.zero_()
    plan.launch()
    ctx.stream.synchronize()
    for seg, r in zip(segs, refs):
        for got, want in zip(seg[6], r):
            assert np.abs(got.numpy() - want).max() <= 1e-14 * max(1.0, np.abs(want).max())
    plan.close()

"""hdgx_local_solve, the element-by-element local solves of kernel/hdg_local.m
(app.getdqdg outside {0, 1}: the elemental step keeps no DUDG_DUH and the
Newton update recomputes each element's dudg, dudg_duh from the state):
  * DUDGT against the oracle's dudg + dudg_duh DUH(:, elcon) (oracle
    elemental, then trace_oracle.local_update) at the FP64 bar 1e-12, for a
    u-only system, a u,q system, 2D p = 4 and tetrahedra p = 2;
  * bit for bit the same as hdgx_elemental + hdgx_local_update on the GPU,
    also when a small workspace limit splits it into chunks;
  * hdg_solve with getdqdg = 2 takes the same Newton steps as with 1."""
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.gpu_util import perturb, relerr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.problems import euler2d_square, euler3d_cube, ns2d_square, with_bdf1  # noqa: E402
from oracle import trace_oracle as TO  # noqa: E402

CASES = {
    "euler2d_p2": lambda: euler2d_square(4, 2, seed=61, alpha=0.3),
    "euler2d_p4": lambda: euler2d_square(3, 4, seed=62),
    "ns2d_p3_bdf1": lambda: with_bdf1(ns2d_square(3, 3, seed=63, av_frac=0.3), 0.05, seed=6),
    "euler3d_p2": lambda: euler3d_cube(2, 2, seed=64, alpha=0.3),
}


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    c = Context(0)
    yield c
    c.set_workspace_limit(0)


def _problem(case):
    P = perturb(CASES[case](), 8)
    rng = np.random.default_rng(9)
    DUH = 1e-2 * rng.standard_normal(P.UH.shape)
    return P, np.asfortranarray(DUH)


@pytest.mark.parametrize("case", list(CASES))
def test_local_solve_matches_oracle_and_local_update(ctx, oracle, case):
    from hdg_amd.runtime import Dev
    P, DUH = _problem(case)
    m = P.mesh
    print(f"[gpu] local solve {case} ne={P.ne}", flush=True)
    args = (P.master, P.app, m.dgnodes, m.bf, P.udg, P.uh, P.sh)
    _, _, DUDG, DUDG_DUH, _ = oracle.elemental(*args)
    ref = TO.local_update(DUDG, DUDG_DUH, DUH, m.elcon)
    got = ctx.local_solve(P.master, P.app, m.dgnodes, m.bf, P.udg, P.uh, P.sh, m.elcon, DUH)
    ctx.stream.synchronize()
    e = relerr(got.numpy(), ref)
    print(f"[gpu]   DUDGT vs oracle {e:.2e}", flush=True)
    assert e < 1e-12
    # the same kernels through hdgx_elemental and hdgx_local_update
    from hdg_amd.trace import TraceSystem
    ts = TraceSystem(ctx, m, P.app.ncu, with_bilu0=False)
    out = ctx.elemental(*args)
    upd = ts.local_update(P.master.npv, P.app.nc, out[2], out[3], Dev.from_numpy(DUH))
    ctx.stream.synchronize()
    assert np.array_equal(got.numpy(), upd.numpy())


def test_local_solve_in_chunks(ctx):
    P, DUH = _problem("euler3d_p2")
    m = P.mesh
    print(f"[gpu] local solve in chunks ne={P.ne}", flush=True)
    args = (P.master, P.app, m.dgnodes, m.bf, P.udg, P.uh, P.sh, m.elcon, DUH)
    one = ctx.local_solve(*args).numpy()
    try:
        ctx.set_workspace_limit(8 << 20)        # ~400 KB per element with its outputs: chunks of 21, 21, 6
        many = ctx.local_solve(*args).numpy()
        ctx.stream.synchronize()
    finally:
        ctx.set_workspace_limit(0)
    assert np.array_equal(one, many)


def test_newton_with_local_solves(ctx):
    """hdg_solve.m's damped Newton with getdqdg = 2 (hdg_local recomputes the
    element solves) against getdqdg = 1 (DUDG_DUH kept): the same steps."""
    import dataclasses
    from hdg_amd.solve import HdgState, hdg_solve
    P = ns2d_square(3, 2, seed=65, av_frac=0.3)
    rng = np.random.default_rng(10)
    udg = np.asfortranarray(P.udg * (1.0 + 1e-2 * rng.standard_normal(P.udg.shape)))
    UH = np.asfortranarray(P.UH * (1.0 + 1e-2 * rng.standard_normal(P.UH.shape)))
    res = {}
    for g in (1, 2):
        print(f"[gpu] Newton NS p2 getdqdg={g}", flush=True)
        app = dataclasses.replace(P.app, getdqdg=g)
        S = HdgState(ctx, P.master, app, P.mesh, udg, UH, P.sh)
        stats = hdg_solve(S, newton_tol=1e-10, gmres_tol=1e-13, maxiter=3)
        res[g] = (S.host(), stats.residuals)
    (u1, h1), r1 = res[1]
    (u2, h2), r2 = res[2]
    assert np.allclose(r1, r2, rtol=1e-9)
    assert relerr(u2, u1) < 1e-12 and relerr(h2, h1) < 1e-12

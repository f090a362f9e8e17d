"""GPU tests of the getdqdg = 0 path and the BDF / DIRK drivers on the real
HDG system:

* hdgx_elemental with getdqdg = 0 (or 2) writes the first ncu (or no) rows of
  DUDG / DUDG_DUH, bit for bit those of getdqdg = 1, and the same AE, FE;
* the Newton solve with getdqdg = 0 (q from getq after every step) converges
  to the getdqdg = 1 solution, and its q columns are getq of its (u, uh);
* a uniform free stream is a fixed point of a DIRK step;
* the observed temporal order of BDF1, DIRK(2,2) and DIRK(3,3) on a Navier-
  Stokes run (errors against a DIRK(3,4) run with an eighth of the step).
"""
import dataclasses

import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.gpu_util import relerr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.problems import euler2d_square, gather_uh, ns2d_square, with_bdf1  # noqa: E402
from oracle import getq_oracle as G  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


def _perturbed(P, seed, amp):
    rng = np.random.default_rng(seed)
    udg = np.asfortranarray(P.udg * (1.0 + amp * rng.standard_normal(P.udg.shape)))
    UH = np.asfortranarray(P.UH * (1.0 + amp * rng.standard_normal(P.UH.shape)))
    return udg, UH


@pytest.mark.parametrize("prec", [0, 1])
def test_elemental_getdqdg(ctx, prec):
    P = ns2d_square(3, 3, seed=51, av_frac=0.3)
    udg, UH = _perturbed(P, 1, 1e-2)
    uh = gather_uh(UH, P.mesh.elcon)
    ncu = P.app.ncu
    print(f"[gpu] elemental getdqdg prec={prec}", flush=True)
    ctx.set_precision(prec)
    try:
        full = ctx.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, udg, uh, P.sh)
        u_only = ctx.elemental(P.master, dataclasses.replace(P.app, getdqdg=0), P.mesh.dgnodes, P.mesh.bf, udg, uh, P.sh)
        none = ctx.elemental(P.master, dataclasses.replace(P.app, getdqdg=2), P.mesh.dgnodes, P.mesh.bf, udg, uh, P.sh)
        ctx.stream.synchronize()
    finally:
        ctx.set_precision(0)
    AE, FE, DUDG, DUDG_DUH, st = (a.numpy() for a in full)
    assert u_only[2].shape == (P.master.npv, ncu, P.ne)
    assert np.array_equal(u_only[2].numpy(), DUDG[:, :ncu])
    assert np.array_equal(u_only[3].numpy(), DUDG_DUH[:, :ncu])
    for got in (u_only, none):
        assert np.array_equal(got[0].numpy(), AE) and np.array_equal(got[1].numpy(), FE)
    assert none[2].shape[1] == 0
    assert not st.any()


def test_newton_getdqdg0(ctx):
    from hdg_amd.solve import HdgState, hdg_solve
    # one implicit Euler step (the steady problem from this state is not a
    # Newton-solvable one)
    P = with_bdf1(ns2d_square(3, 2, seed=52, av_frac=0.3), 0.02, seed=5)
    udg, UH = _perturbed(P, 2, 1e-2)
    print("[gpu] Newton getdqdg 0 vs 1, ns2d p2 BDF1", flush=True)
    S1 = HdgState(ctx, P.master, P.app, P.mesh, udg, UH, P.sh)
    st1 = hdg_solve(S1, newton_tol=1e-11, gmres_tol=1e-13)
    S0 = HdgState(ctx, P.master, dataclasses.replace(P.app, getdqdg=0), P.mesh, udg, UH, P.sh)
    st0 = hdg_solve(S0, newton_tol=1e-11, gmres_tol=1e-13)
    print(f"[gpu] residuals getdqdg=1 {st1.residuals}", flush=True)
    print(f"[gpu] residuals getdqdg=0 {st0.residuals}", flush=True)
    assert st0.residuals[-1] <= 1e-11 * st0.residuals[0]
    u0, h0 = S0.host()
    u1, h1 = S1.host()
    assert relerr(h0, h1) < 1e-9 and relerr(u0, u1) < 1e-9
    ncu = P.app.ncu
    q = G.getq(P.master, P.mesh.dgnodes, u0, gather_uh(h0, P.mesh.elcon), fc_q=P.app.fc_q)
    assert relerr(u0[:, ncu:], q) < 1e-12


def test_dirk_free_stream_is_fixed(ctx):
    from hdg_amd.solve import HdgState
    from hdg_amd.timestep import hdg_solve_dirk
    P = euler2d_square(3, 2, seed=53, alpha=0.3)
    ncu = P.app.ncu
    uinf = np.asarray(P.app.bcs[0][:ncu])
    udg = np.zeros_like(P.udg)
    udg[:, :ncu, :] = uinf[None, :, None]
    UH = np.asfortranarray(np.repeat(uinf[:, None], P.UH.shape[1], axis=1))
    print("[gpu] DIRK(3,3) on a free stream", flush=True)
    S = HdgState(ctx, P.master, dataclasses.replace(P.app), P.mesh, udg, UH, None)
    hdg_solve_dirk(S, 0.0, 0.05, 3, 3, newton_tol=1e-12, gmres_tol=1e-13)
    u, h = S.host()
    assert relerr(u, udg) < 1e-11 and relerr(h, UH) < 1e-11


def _integrate(ctx, P, udg, UH, method, dt, T):
    from hdg_amd.solve import HdgState
    from hdg_amd.timestep import hdg_solve_bdf, hdg_solve_dirk
    S = HdgState(ctx, P.master, dataclasses.replace(P.app), P.mesh, udg, UH, None)
    nsteps = int(round(T / dt))
    hist = [S.UDG.t.clone()]
    worst = 0.0
    for n in range(nsteps):
        if method[0] == "bdf":
            st, hist = hdg_solve_bdf(S, hist, (n + 1) * dt, dt, method[1], newton_tol=1e-11, gmres_tol=1e-13)
            st = [st]
        else:
            st = hdg_solve_dirk(S, n * dt, dt, method[1], method[2], newton_tol=1e-11, gmres_tol=1e-13)
        for s in st:
            # converged: relative 1e-11, or down at the round-off floor
            worst = max(worst, s.residuals[-1] / (1e-11 * s.residuals[0] + 1e-12))
    assert worst <= 1.0, (method, dt, worst)
    return S.host()


def test_temporal_order(ctx):
    """The error against a DIRK(3,4) run at dt/32 falls by 2^p from dt/2 to
    dt/4 for a method of order p (BDF1 is still slightly pre-asymptotic: the
    acoustic modes are stiff at these steps); the self-convergence rate
    log2(|u(dt) - u(dt/2)| / |u(dt/2) - u(dt/4)|) is printed beside it."""
    # the smooth initial state of the problem (random noise would start a
    # stiff transient that hides the asymptotic rates)
    P = ns2d_square(2, 2, seed=54, av_frac=0.3)
    udg, UH = P.udg, P.UH
    T, dt0 = 0.02, 0.01
    ncu = P.app.ncu
    print("[gpu] temporal order: reference DIRK(3,4)", flush=True)
    ref_u = _integrate(ctx, P, udg, UH, ("dirk", 3, 4), dt0 / 32, T)[0][:, :ncu]
    for method, p in ((("bdf", 1), 1), (("dirk", 2, 2), 2), (("dirk", 3, 3), 3)):
        us = [_integrate(ctx, P, udg, UH, method, dt0 / 2 ** k, T)[0][:, :ncu] for k in range(3)]
        d01, d12 = np.abs(us[0] - us[1]).max(), np.abs(us[1] - us[2]).max()
        errs = [np.abs(u - ref_u).max() for u in us]
        rate = np.log2(errs[1] / errs[2])
        print(f"[gpu] {method}: order {rate:.2f} (self-convergence {np.log2(d01 / d12):.2f}); errors {errs}",
              flush=True)
        assert abs(rate - p) < 0.25, (method, rate, errs)

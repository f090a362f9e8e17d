"""GPU parity of the trace solve and Newton update (the hdgx_trace_*,
hdgx_bilu0_*, hdgx_gmres, hdgx_local_update, hdgx_newton_update entry points)
against the oracle of oracle/trace_oracle.py on the same inputs.

Bars: FP64 block system, operator, ILU(0) and local update <= 1e-12
normwise relative; GMRES at tol 1e-12 against a direct sparse solve <= 1e-9;
gathers and axpy-type updates exact to round-off."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests.conftest import has_gpu
from tests.gpu_util import relerr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.problems import euler2d_square, euler3d_cube, gather_uh, ns2d_square  # noqa: E402
from oracle import trace_oracle as TO  # noqa: E402

CASES = {
    "euler2d_p1": lambda: euler2d_square(4, 1, seed=41, alpha=0.3),
    "euler2d_p2": lambda: euler2d_square(4, 2, seed=42, alpha=0.3),
    "euler2d_p4": lambda: euler2d_square(3, 4, seed=44, alpha=0.3),
    "ns2d_p3": lambda: ns2d_square(3, 3, seed=45, av_frac=0.3),
    "euler3d_p1": lambda: euler3d_cube(2, 1, seed=46, alpha=0.3),
}


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


def _system(oracle, case, seed=3, amp=1e-3):
    P = CASES[case]()
    rng = np.random.default_rng(seed)
    udg = np.asfortranarray(P.udg * (1.0 + amp * rng.standard_normal(P.udg.shape)))
    UH = np.asfortranarray(P.UH * (1.0 + amp * rng.standard_normal(P.UH.shape)))
    uh = gather_uh(UH, P.mesh.elcon)
    AE, FE, DUDG, DUDG_DUH, st = oracle.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, udg, uh, P.sh)
    assert not st.any()
    return P, udg, UH, AE, FE, DUDG, DUDG_DUH


def _direct(P, AE, FE):
    A, b = TO.sparse_system(AE, FE, P.mesh.elcon, P.app.ncu, P.mesh.nsiz)
    return spla.spsolve(A.tocsc(), b)


@pytest.mark.parametrize("case", list(CASES))
def test_trace_assemble_and_matvec(ctx, oracle, case):
    from hdg_amd.runtime import Dev
    from hdg_amd.trace import TraceSystem
    P, udg, UH, AE, FE, *_ = _system(oracle, case)
    print(f"[gpu] trace assemble {case} nf={P.mesh.nf}", flush=True)
    ts = TraceSystem(ctx, P.mesh, P.app.ncu)
    K, F, st = ts.assemble(Dev.from_numpy(AE), Dev.from_numpy(FE))
    x = np.random.default_rng(5).standard_normal((ts.ncf, ts.nf))
    y = ts.matvec(K, Dev.from_numpy(x))
    ctx.stream.synchronize()
    assert not st.numpy().any()
    Kg, Fg, yg = K.numpy(), F.numpy(), y.numpy()
    if P.mesh.nd == 2:
        # face permutations are involutions in 2D: the reference's indexing applies as written
        Kr, Fr = TO.mk_dense_block_system(AE, FE, P.mesh.f, P.mesh.t2f, P.mesh.elcon)
        assert relerr(Kg, Kr) < 1e-12 and relerr(Fg, Fr) < 1e-12, (relerr(Kg, Kr), relerr(Fg, Fr))
        yr = TO.block_matvec(Kr, ts.f2f_host, x)
        assert relerr(yg, yr) < 1e-12
    # in every dimension the block system's solution is the global system's
    xd = _direct(P, AE, FE)
    ncf, nf = ts.ncf, ts.nf
    A = np.eye(ncf * nf)
    for i in range(nf):
        for k in range(ts.nbf - 1):
            j = ts.f2f_host[k + 1, i]
            if j >= 0:
                A[i * ncf:(i + 1) * ncf, j * ncf:(j + 1) * ncf] += Kg[:, :, k, i]
    xb = np.linalg.solve(A, Fg.ravel(order="F"))
    assert relerr(xb, xd) < 1e-10


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precond", ["none", "bilu0"])
def test_gmres_solves_the_trace_system(ctx, oracle, case, precond):
    from hdg_amd.runtime import Dev
    from hdg_amd.trace import TraceSystem
    P, udg, UH, AE, FE, *_ = _system(oracle, case)
    print(f"[gpu] gmres {case} {precond} nf={P.mesh.nf}", flush=True)
    ts = TraceSystem(ctx, P.mesh, P.app.ncu)
    K, F, st = ts.assemble(Dev.from_numpy(AE), Dev.from_numpy(FE))
    LU = None
    if precond == "bilu0":
        LU, lst = ts.bilu0_factor(K)
    x, it, rr, fl = ts.gmres(K, F, LU=LU, restart=20, tol=1e-12, maxit=2000)
    xg = x.numpy()
    xd = _direct(P, AE, FE)
    print(f"[gpu]   iters={it} relres={rr:.2e} flag={fl} err={relerr(xg.ravel(order='F'), xd):.2e}", flush=True)
    assert fl == 0 and rr <= 1e-12
    assert relerr(xg.ravel(order="F"), xd) < 1e-9
    if precond == "none":
        # same algorithm as kernel/gmres.m: iteration counts agree
        shape = xg.shape
        Kg = K.numpy()
        op = lambda v: TO.block_matvec(Kg, ts.f2f_host, v.reshape(shape, order="F")).ravel(order="F")
        _, flag, it_ref, _ = TO.gmres(op, F.numpy().ravel(order="F"), restart=20, tol=1e-12, maxit=2000)
        assert abs(it - it_ref) <= 2, (it, it_ref)


@pytest.mark.parametrize("ordering", ["colors", "natural"])
@pytest.mark.parametrize("case", ["euler2d_p2", "ns2d_p3", "euler3d_p1"])
def test_bilu0_factor_and_apply(ctx, oracle, case, ordering):
    """Block ILU(0) factor and apply against computeBILU0/applyBILU0 in the
    same row order: the greedy face colouring (3 colours on triangles, more on
    tetrahedra) or one colour per face (the natural order, every sweep one
    face: the apply's first-colour, middle-colour and last-colour paths)."""
    from hdg_amd.runtime import Dev
    from hdg_amd.trace import TraceSystem, face_connectivity, mkf2f
    P, udg, UH, AE, FE, *_ = _system(oracle, case)
    print(f"[gpu] bilu0 {case} {ordering}", flush=True)
    conn = None
    if ordering == "natural":
        f2f = mkf2f(P.mesh)
        nf = f2f.shape[1]
        ids = np.arange(nf, dtype=np.int32)
        conn = dict(fconn=face_connectivity(P.mesh), f2f=f2f,
                    color=(np.arange(nf + 1, dtype=np.int32), ids, ids, ids))
    ts = TraceSystem(ctx, P.mesh, P.app.ncu, conn=conn)
    if ordering == "natural":
        assert ts.ncolors == ts.nf
    K, F, st = ts.assemble(Dev.from_numpy(AE), Dev.from_numpy(FE))
    LU, lst = ts.bilu0_factor(K)
    x = np.random.default_rng(8).standard_normal((ts.ncf, ts.nf))
    y = ts.bilu0_apply(LU, Dev.from_numpy(x))
    yg = y.numpy()
    assert not lst.numpy().any()
    # oracle: computeBILU0 / applyBILU0 on the same rows in the same order
    Kg = K.numpy()
    rowpts, colind, blocks = [0], [], []
    for i in range(ts.nf):
        colind.append(i)
        blocks.append(np.eye(ts.ncf))
        for k in range(ts.nbf - 1):
            j = ts.f2f_host[k + 1, i]
            if j >= 0:
                colind.append(j)
                blocks.append(Kg[:, :, k, i])
        rowpts.append(len(colind))
    rowpts, colind, Hg = np.array(rowpts), np.array(colind), np.stack(blocks, axis=2)
    order = ts.color_rows.numpy()
    Mg = TO.bilu0(Hg, rowpts, colind, order)
    yr = TO.bilu0_apply(Mg, rowpts, colind, order, x)
    assert relerr(yg, yr) < 1e-12, relerr(yg, yr)


@pytest.mark.parametrize("case", ["euler2d_p2", "ns2d_p3", "euler3d_p1"])
def test_trace_scatter(ctx, oracle, case):
    """hdgx_trace_scatter: the per-element Rh summed at the global trace
    unknowns (hdg_residual.m:41-52), against a numpy scatter-add over elcon.
    Every trace unknown has one or two terms, so the sums agree to the bit."""
    from hdg_amd.runtime import Dev
    from hdg_amd.trace import TraceSystem
    P = CASES[case]()
    print(f"[gpu] trace scatter {case}", flush=True)
    ts = TraceSystem(ctx, P.mesh, P.app.ncu, with_bilu0=False)
    ndf = ts.npf * ts.nfe
    Rhe = np.asfortranarray(np.random.default_rng(5).standard_normal((ts.ncu, ndf, ts.ne)))
    RH = ts.scatter(Dev.from_numpy(Rhe)).numpy().reshape(ts.ncu, ts.npf * ts.nf, order="F")
    ref = np.zeros((ts.ncu, ts.npf * ts.nf))
    for e in range(ts.ne):
        for a in range(ndf):
            ref[:, P.mesh.elcon[a, e]] += Rhe[:, a, e]
    assert np.array_equal(RH, ref)


@pytest.mark.parametrize("case", ["euler2d_p2", "ns2d_p3", "euler3d_p1"])
def test_gather_local_and_newton_update(ctx, oracle, case):
    from hdg_amd.runtime import Dev
    from hdg_amd.trace import TraceSystem
    P, udg, UH, AE, FE, DUDG, DUDG_DUH = _system(oracle, case)
    print(f"[gpu] local update {case}", flush=True)
    ts = TraceSystem(ctx, P.mesh, P.app.ncu, with_bilu0=False)
    UHd = Dev.from_numpy(UH)
    uhe = ts.gather(UHd).numpy()
    assert np.array_equal(uhe, gather_uh(UH, P.mesh.elcon))
    DUH = _direct(P, AE, FE).reshape(P.app.ncu, -1, order="F")
    npv, nco = DUDG.shape[0], DUDG.shape[1]
    DUDGT = ts.local_update(npv, nco, Dev.from_numpy(DUDG), Dev.from_numpy(DUDG_DUH), Dev.from_numpy(DUH))
    ref = TO.local_update(DUDG, DUDG_DUH, DUH, P.mesh.elcon)
    assert relerr(DUDGT.numpy(), ref) < 1e-12
    U = Dev.from_numpy(udg)
    UH2 = Dev.from_numpy(UH)
    ts.newton_update(npv, udg.shape[1], nco, 0.5, U, DUDGT, UH2, Dev.from_numpy(DUH), U, UH2)
    exp_u = udg.copy()
    exp_u[:, :nco, :] += 0.5 * ref
    assert relerr(U.numpy(), exp_u) < 1e-14
    assert relerr(UH2.numpy(), UH + 0.5 * DUH) < 1e-15
    assert abs(ts.norm2(Dev.from_numpy(DUH)) - np.linalg.norm(DUH)) <= 1e-13 * np.linalg.norm(DUH)


def test_device_newton_matches_the_oracle_newton(ctx, oracle):
    """hdg_solve on the device vs the oracle's Newton (direct trace solves):
    same residual history while the residual is far above the GMRES tolerance,
    and the same converged state."""
    from hdg_amd.solve import HdgState, hdg_solve
    P = euler2d_square(3, 2, seed=9, alpha=0.3)
    rng = np.random.default_rng(4)
    udg = np.asfortranarray(P.udg * (1.0 + 2e-2 * rng.standard_normal(P.udg.shape)))
    UH = np.asfortranarray(P.UH * (1.0 + 2e-2 * rng.standard_normal(P.UH.shape)))
    print("[gpu] device Newton euler2d p2", flush=True)
    S = HdgState(ctx, P.master, P.app, P.mesh, udg, UH, P.sh)
    stats = hdg_solve(S, newton_tol=1e-10, gmres_tol=1e-13, maxiter=8)
    ug, uhg = S.host()
    # oracle Newton with direct solves, undamped (the device run accepts alpha = 1 here)
    assert all(a == 1.0 for a in stats.alphas)
    m = P.mesh
    u, H = udg.copy(), UH.copy()
    res = []
    for _ in range(len(stats.alphas) + 1):
        uh = gather_uh(H, m.elcon)
        AE, FE, DUDG, DUDG_DUH, _ = oracle.elemental(P.master, P.app, m.dgnodes, m.bf, u, uh, P.sh)
        K, F = TO.mk_dense_block_system(AE, FE, m.f, m.t2f, m.elcon)
        res.append(np.linalg.norm(F))
        DUH = _direct(P, AE, FE).reshape(P.app.ncu, -1, order="F")
        u[:, :DUDG.shape[1], :] += TO.local_update(DUDG, DUDG_DUH, DUH, m.elcon)
        H = H + DUH
    r_dev = np.array(stats.residuals)
    r_ref = np.array(res[:len(r_dev)])
    big = r_ref > 1e-6 * r_ref[0]
    assert np.allclose(r_dev[big], r_ref[big], rtol=1e-6), (r_dev, r_ref)
    assert r_dev[-1] <= 1e-10 * r_dev[0]
    e_h, e_u = relerr(uhg, H), relerr(ug, u)
    print(f"[gpu]   converged state vs oracle: UH {e_h:.2e} UDG {e_u:.2e}", flush=True)
    assert e_h < 1e-12 and e_u < 1e-12

"""Pins the CPU restatement of the trace solve and Newton update
(oracle/trace_oracle.py) by checks that do not restate it: direct solves of
the sparse global system, ILU(0) identities, and Newton convergence."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from hdg_amd.problems import euler2d_square, gather_uh, ns2d_square
from oracle import trace_oracle as TO


def _perturb(P, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    udg = P.udg * (1.0 + amp * rng.standard_normal(P.udg.shape))
    UH = P.UH * (1.0 + amp * rng.standard_normal(P.UH.shape))
    return np.asfortranarray(udg), np.asfortranarray(UH)


@pytest.fixture(scope="module")
def euler_sys(oracle):
    P = euler2d_square(3, 2, seed=5, alpha=0.3)
    udg, UH = _perturb(P, 11)
    uh = gather_uh(UH, P.mesh.elcon)
    AE, FE, DUDG, DUDG_DUH, st = oracle.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, udg, uh, P.sh)
    assert not st.any()
    return P, AE, FE, DUDG, DUDG_DUH


def _block_dense(K, f2f):
    """Face-block system as one dense matrix (identity diagonal blocks)."""
    ncf, _, nb1, nf = K.shape
    A = np.eye(ncf * nf)
    for i in range(nf):
        for k in range(nb1):
            j = f2f[k + 1, i]
            if j >= 0:
                A[i * ncf:(i + 1) * ncf, j * ncf:(j + 1) * ncf] += K[:, :, k, i]
    return A


def test_block_system_solves_the_global_system(euler_sys):
    from hdg_amd.trace import mkf2f
    P, AE, FE, *_ = euler_sys
    m = P.mesh
    ncu = P.app.ncu
    A, b = TO.sparse_system(AE, FE, m.elcon, ncu, m.nsiz)
    x_sparse = spla.spsolve(A.tocsc(), b)
    K, F = TO.mk_dense_block_system(AE, FE, m.f, m.t2f, m.elcon)
    f2f = mkf2f(m)
    x_block = np.linalg.solve(_block_dense(K, f2f), F.ravel(order="F"))
    assert np.linalg.norm(x_block - x_sparse) <= 1e-10 * np.linalg.norm(x_sparse)
    # and the block operator is the dense one
    v = np.random.default_rng(0).standard_normal(F.shape)
    y = TO.block_matvec(K, f2f, v)
    assert np.allclose(y.ravel(order="F"), _block_dense(K, f2f) @ v.ravel(order="F"), rtol=1e-13, atol=1e-13)


def test_gmres_converges_to_the_direct_solution(euler_sys):
    from hdg_amd.trace import mkf2f
    P, AE, FE, *_ = euler_sys
    m = P.mesh
    K, F = TO.mk_dense_block_system(AE, FE, m.f, m.t2f, m.elcon)
    f2f = mkf2f(m)
    shape = F.shape
    op = lambda v: TO.block_matvec(K, f2f, v.reshape(shape, order="F")).ravel(order="F")
    x, flag, it, rev = TO.gmres(op, F.ravel(order="F"), restart=20, tol=1e-12)
    xd = np.linalg.solve(_block_dense(K, f2f), F.ravel(order="F"))
    assert flag == 0
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    assert np.all(np.diff(rev[:5]) <= 1e-15 + 0 * rev[:4]) or rev[-1] <= 1e-12


def _cbsr_from_f2f(K, f2f):
    """CBSR of the scaled face-block system: row i = [i, its f2f neighbours]."""
    ncf, _, nb1, nf = K.shape
    rowpts, colind, blocks = [0], [], []
    for i in range(nf):
        colind.append(i)
        blocks.append(np.eye(ncf))
        for k in range(nb1):
            j = f2f[k + 1, i]
            if j >= 0:
                colind.append(j)
                blocks.append(K[:, :, k, i])
        rowpts.append(len(colind))
    return np.array(rowpts), np.array(colind), np.stack(blocks, axis=2)


def _lu_blocks(Mg, rowpts, colind, order):
    """Dense L (unit lower) and U from the ILU(0) storage."""
    nrows = len(rowpts) - 1
    bsz = Mg.shape[0]
    rank = np.empty(nrows, dtype=int)
    rank[order] = np.arange(nrows)
    L = np.eye(bsz * nrows)
    U = np.zeros((bsz * nrows, bsz * nrows))
    for i in range(nrows):
        si = slice(i * bsz, (i + 1) * bsz)
        U[si, si] = np.linalg.inv(Mg[:, :, rowpts[i]])
        for k in range(rowpts[i] + 1, rowpts[i + 1]):
            j = colind[k]
            sj = slice(j * bsz, (j + 1) * bsz)
            if rank[j] < rank[i]:
                L[si, sj] = Mg[:, :, k]
            else:
                U[si, sj] = Mg[:, :, k]
    return L, U


@pytest.mark.parametrize("order_kind", ["natural", "reversed"])
def test_bilu0_matches_the_matrix_on_its_pattern(euler_sys, order_kind):
    P, AE, FE, *_ = euler_sys
    m = P.mesh
    rowpts, colind = TO.entcon(m.t2f)
    Hg, Rg = TO.global_assembly(AE, FE, m.t2f, m.elcon, rowpts, colind, P.app.ncu)
    nrows = len(rowpts) - 1
    order = np.arange(nrows) if order_kind == "natural" else np.arange(nrows)[::-1].copy()
    Mg = TO.bilu0(Hg, rowpts, colind, order)
    L, U = _lu_blocks(Mg, rowpts, colind, order)
    # permuted: rows/cols in elimination order make L lower and U upper
    bsz = Hg.shape[0]
    LU = L @ U
    for i in range(nrows):
        for k in range(rowpts[i], rowpts[i + 1]):
            j = colind[k]
            blk = LU[i * bsz:(i + 1) * bsz, j * bsz:(j + 1) * bsz]
            assert np.allclose(blk, Hg[:, :, k], rtol=1e-10, atol=1e-10 * np.abs(Hg).max())
    # the apply is the solve with L U
    x = np.random.default_rng(1).standard_normal(Rg.shape)
    v = TO.bilu0_apply(Mg, rowpts, colind, order, x)
    assert np.allclose((L @ U) @ v.ravel(order="F"), x.ravel(order="F"), rtol=1e-9, atol=1e-9)


def test_bilu0_is_exact_on_a_block_tridiagonal_matrix():
    rng = np.random.default_rng(3)
    n, b = 7, 3
    rowpts, colind, blocks = [0], [], []
    for i in range(n):
        nb = [i] + [j for j in (i - 1, i + 1) if 0 <= j < n]
        for j in nb:
            colind.append(j)
            blocks.append(rng.standard_normal((b, b)) + (4 * np.eye(b) if j == i else 0))
        rowpts.append(len(colind))
    rowpts, colind, Hg = np.array(rowpts), np.array(colind), np.stack(blocks, axis=2)
    Mg = TO.bilu0(Hg, rowpts, colind, np.arange(n))
    x = rng.standard_normal((b, n))
    v = TO.bilu0_apply(Mg, rowpts, colind, np.arange(n), x)
    y = TO.cbsr_matvec(Hg, rowpts, colind, v)
    assert np.allclose(y, x, rtol=1e-12, atol=1e-12)


def test_bilu0_of_the_scaled_system_is_the_same_preconditioner(euler_sys):
    """Block ILU(0) commutes with block-row scaling: the preconditioned
    operator of the scaled face-block system equals that of the assembled one."""
    from hdg_amd.trace import color_faces, mkf2f
    P, AE, FE, *_ = euler_sys
    m = P.mesh
    rowpts, colind = TO.entcon(m.t2f)
    Hg, Rg = TO.global_assembly(AE, FE, m.t2f, m.elcon, rowpts, colind, P.app.ncu)
    K, F = TO.mk_dense_block_system(AE, FE, m.f, m.t2f, m.elcon)
    f2f = mkf2f(m)
    rp2, ci2, Kg = _cbsr_from_f2f(K, f2f)
    _, rows, rank, _ = color_faces(f2f)
    order = rows
    M1 = TO.bilu0(Hg, rowpts, colind, order)
    M2 = TO.bilu0(Kg, rp2, ci2, order)
    x = np.random.default_rng(2).standard_normal(Rg.shape)
    a = TO.bilu0_apply(M1, rowpts, colind, order, TO.cbsr_matvec(Hg, rowpts, colind, x))
    b = TO.bilu0_apply(M2, rp2, ci2, order, TO.cbsr_matvec(Kg, rp2, ci2, x))
    assert np.allclose(a, b, rtol=1e-9, atol=1e-9 * np.abs(a).max())
    # and it solves the system with gmres_bilu0
    mv = lambda v: TO.cbsr_matvec(Hg, rowpts, colind, v.reshape(Rg.shape, order="F")).ravel(order="F")
    pc = lambda v: TO.bilu0_apply(M1, rowpts, colind, order, v.reshape(Rg.shape, order="F")).ravel(order="F")
    xs, flag, it = TO.gmres_bilu0(mv, pc, Rg.ravel(order="F"), tol=1e-12)
    A, bb = TO.sparse_system(AE, FE, m.elcon, P.app.ncu, m.nsiz)
    xd = spla.spsolve(A.tocsc(), bb)
    assert flag == 0 and np.linalg.norm(xs - xd) <= 1e-9 * np.linalg.norm(xd)


def _newton(oracle, P, udg, UH, iters=4):
    m = P.mesh
    ncu = P.app.ncu
    norms = []
    for _ in range(iters):
        uh = gather_uh(UH, m.elcon)
        AE, FE, DUDG, DUDG_DUH, st = oracle.elemental(P.master, P.app, m.dgnodes, m.bf, udg, uh, P.sh)
        A, b = TO.sparse_system(AE, FE, m.elcon, ncu, m.nsiz)
        norms.append(np.linalg.norm(b))
        DUH = spla.spsolve(A.tocsc(), b).reshape(ncu, m.nsiz, order="F")
        DUDGT = TO.local_update(DUDG, DUDG_DUH, DUH, m.elcon)
        nco = DUDGT.shape[1]
        udg = udg.copy()
        udg[:, :nco, :] += DUDGT
        UH = UH + DUH
    return norms, udg, UH


def test_newton_converges_quadratically(oracle):
    """hdg_solve.m:105-170 with the direct solve: the residual of a perturbed
    state drops quadratically, which a wrong local update would destroy."""
    P = euler2d_square(2, 2, seed=9, alpha=0.3)
    udg, UH = _perturb(P, 4, amp=2e-2)
    norms, *_ = _newton(oracle, P, udg, UH, iters=4)
    r = np.array(norms)
    assert r[1] < 0.2 * r[0]
    assert r[2] < 0.2 * r[1] and r[3] < max(1e-12, 10 * r[2] ** 2 / r[1] ** 2 * r[2])


def test_uniform_state_gives_a_zero_update(oracle):
    P = euler2d_square(2, 2, seed=0)
    m = P.mesh
    udg = np.zeros_like(P.udg)
    for c in range(P.app.ncu):
        udg[:, c, :] = P.udg[0, c, 0]
    UH = np.repeat(udg[0, :P.app.ncu, 0][:, None], m.nsiz, axis=1)
    # far-field states of the boundary equal the uniform state
    norms, udg2, UH2 = _newton(oracle, P, udg, UH, iters=1)
    assert norms[0] < 1e-10
    assert np.abs(UH2 - UH).max() < 1e-10 and np.abs(udg2 - udg).max() < 1e-10


def _greedy_colours(f2f):
    """The greedy colouring spelled out: face i takes the smallest colour that
    none of its already coloured neighbours has."""
    nbf, nf = f2f.shape
    color = -np.ones(nf, dtype=np.int64)
    for i in range(nf):
        used = {int(color[j]) for j in f2f[1:, i] if j >= 0 and color[j] >= 0}
        c = 0
        while c in used:
            c += 1
        color[i] = c
    return color


@pytest.mark.parametrize("nd,p,n", [(2, 1, 6), (2, 3, 5), (3, 2, 2)])
def test_face_colouring(nd, p, n):
    from hdg_amd.master import mkmaster
    from hdg_amd.mesh import mkmesh_cube, mkmesh_square
    from hdg_amd.runtime import HdgxError
    from hdg_amd.trace import check_coloring, color_faces, mkf2f
    m = mkmaster(nd, p)
    mesh = mkmesh_square(n, n, m, jitter_amp=0.1, seed=n) if nd == 2 else mkmesh_cube(n, m, jitter_amp=0.1, seed=n)
    f2f = mkf2f(mesh)
    ptr, rows, rank, color = color_faces(f2f)
    assert np.array_equal(color, _greedy_colours(f2f))
    assert check_coloring(f2f, color)
    assert ptr[-1] == f2f.shape[1] and np.array_equal(rows[rank], np.arange(f2f.shape[1]))
    assert np.all(np.diff(color[rows]) >= 0)
    bad = f2f.copy()
    bad[1, 0] = f2f.shape[1]
    with pytest.raises(HdgxError):
        color_faces(bad)

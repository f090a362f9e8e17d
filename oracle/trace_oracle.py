"""CPU restatement of the reference's trace solve and Newton update, for tests.

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the product path
(hdg_amd/).  FP64 numpy, written for small meshes (Python loops over faces).

Functions and the reference code each follows:
  mk_dense_block_system  kernel/mkDenseBlockSystem.m:9-95 (mkBlockSystem1)
  sparse_system          kernel/hdg_assemble.m:76-87 (the sparse branch)
  block_matvec           kernel/gmres.m:38-44 with kernel/arraymul.c:53-70
  gmres                  kernel/gmres.m:1-109 (CGS, least squares per step)
  entcon                 mesh/mkmesh/elcon2entcon.m:1-45 for HDG (t2f')
  global_assembly        kernel/globalassembly.m:1-39 (hdg branch)
  bilu0 / bilu0_apply    kernel/computeBILU0.m:1-45, kernel/applyBILU0.m:1-52
  cbsr_matvec            kernel/matvec.m:1-12
  gmres_bilu0            kernel/gmres_bilu0.m:1-103 (MGS)
  local_update           kernel/hdg_solve.m:119-121 (mapContractK)
All indices are 0-based.  The parity of each function is pinned in
tests/test_trace_oracle.py by checks that do not restate it: the dense block
system against a direct solve of the sparse global system, GMRES against that
direct solve, ILU(0) against exactness on a block tridiagonal system and the
ILU(0) pattern identity (LU)_ij = A_ij, the local update against the
element's full linear system.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _face_local(t2f, e, i):
    return int(np.nonzero(t2f[e] == i)[0][0])


def mk_dense_block_system(AE, FE, f, t2f, elcon):
    """K (ncf, ncf, nbf-1, nf), F (ncf, nf): face-block rows scaled by the
    inverse diagonal block (mkDenseBlockSystem.m:37-93)."""
    nch, ndf = AE.shape[0], AE.shape[1]
    ne = AE.shape[4]
    nf = f.shape[0]
    nfe = t2f.shape[1]
    npf = ndf // nfe
    ncf = nch * npf
    nbf = 2 * nfe - 1
    FE = FE.reshape((nch, npf, nfe, ne), order="F")
    AE = AE.reshape((nch, npf, nfe, nch, npf, nfe, ne), order="F")
    el = elcon.reshape((npf, nfe, ne), order="F")
    K = np.zeros((ncf, ncf, nbf - 1, nf))
    F = np.zeros((ncf, nf))

    def blk(e, a_idx, ia, b_idx, ib):
        sub = AE[:, :, ia, :, :, ib, e]                      # (nch, npf, nch, npf)
        return sub[:, a_idx][:, :, :, b_idx].reshape((ncf, ncf), order="F")

    for i in range(nf):
        e1, e2 = int(f[i, -2]), int(f[i, -1])
        i1 = _face_local(t2f, e1, i)
        j1 = el[:, i1, e1] - i * npf
        M = blk(e1, j1, i1, j1, i1)
        Fi = FE[:, :, i1, e1][:, j1].reshape(ncf, order="F")
        k = 0
        for is_ in range(nfe):
            if is_ != i1:
                j = t2f[e1, is_]
                j3 = el[:, is_, e1] - j * npf
                K[:, :, k, i] = blk(e1, j1, i1, j3, is_)
                k += 1
        if e2 >= 0:
            i2 = _face_local(t2f, e2, i)
            j2 = el[:, i2, e2] - i * npf
            M = M + blk(e2, j2, i2, j2, i2)
            Fi = Fi + FE[:, :, i2, e2][:, j2].reshape(ncf, order="F")
            for is_ in range(nfe):
                if is_ != i2:
                    j = t2f[e2, is_]
                    j4 = el[:, is_, e2] - j * npf
                    K[:, :, k, i] = blk(e2, j2, i2, j4, is_)
                    k += 1
        Mi = np.linalg.inv(M)
        F[:, i] = Mi @ Fi
        for kk in range(nbf - 1):
            K[:, :, kk, i] = Mi @ K[:, :, kk, i]
    return K, F


def sparse_system(AE, FE, elcon, ncu, nsiz):
    """Global trace system assembled entry by entry (hdg_assemble.m:77-87):
    unknown (c, node) -> c + ncu*node."""
    ndf, ne = elcon.shape
    nb = ncu * ndf
    con = (elcon[None, :, :] * ncu + np.arange(ncu)[:, None, None]).reshape(nb, ne, order="F")
    il = np.repeat(con[:, None, :], nb, axis=1)
    jl = np.repeat(con[None, :, :], nb, axis=0)
    A = sp.coo_matrix((AE.reshape(nb * nb * ne, order="F"),
                       (il.reshape(-1, order="F"), jl.reshape(-1, order="F"))), shape=(ncu * nsiz, ncu * nsiz)).tocsr()
    b = np.zeros(ncu * nsiz)
    np.add.at(b, con.reshape(-1, order="F"), FE.reshape(-1, order="F"))
    return A, b


def block_matvec(K, f2f, x):
    """y = x + sum_k K_k x(f2f(k)) (gmres.m:38-44); x (ncf, nf)."""
    ncf, _, nb1, nf = K.shape
    y = x.copy()
    for i in range(nf):
        acc = np.zeros(ncf)
        for k in range(nb1):
            j = f2f[k + 1, i]
            if j >= 0:
                acc += K[:, :, k, i] @ x[:, j]
        y[:, i] += acc
    return y


def gmres(op, b, x=None, restart=20, tol=1e-6, maxit=None):
    """kernel/gmres.m:1-109 with op(v) the block operator: classical
    Gram-Schmidt (:81-82), least-squares solve and residual each step (:91-94).
    Returns (x, flag, iter, rev)."""
    N = b.size
    x = np.zeros(N) if x is None else x.copy()
    maxit = N if maxit is None else maxit
    nrmb = np.linalg.norm(b)
    flags, it = np.inf, 0
    rev = []
    y = np.zeros(0)
    while True:
        r = b - op(x)
        beta = np.linalg.norm(r)
        V = np.zeros((N, restart + 1))
        hh = np.zeros((restart + 1, restart))
        V[:, 0] = r / beta
        res = beta
        it += 1
        rev.append(res)
        y = np.zeros(0)
        for j in range(restart):
            if res / nrmb <= tol:
                flags = 0
                break
            if it >= maxit:
                flags = 1
                break
            w = op(V[:, j])
            hh[:j + 1, j] = V[:, :j + 1].T @ w
            w = w - V[:, :j + 1] @ hh[:j + 1, j]
            hh[j + 1, j] = np.linalg.norm(w)
            if hh[j + 1, j] != 0.0:
                V[:, j + 1] = w / hh[j + 1, j]
            else:
                break
            e1 = np.zeros(j + 2)
            e1[0] = beta
            y = np.linalg.lstsq(hh[:j + 2, :j + 1], e1, rcond=None)[0]
            res = np.linalg.norm(e1 - hh[:j + 2, :j + 1] @ y)
            it += 1
            rev.append(res)
        x = x + V[:, :y.size] @ y
        if flags < np.inf:
            break
    return x, flags, it, np.array(rev) / nrmb


def entcon(t2f):
    """Face-to-face CBSR: rows = faces, row i lists i first, then the other
    faces sharing an element with it in ascending order (elcon2entcon.m:27-45
    applied to t2f')."""
    ne, nfe = t2f.shape
    nf = int(t2f.max()) + 1
    elems = [[] for _ in range(nf)]
    for e in range(ne):
        for j in range(nfe):
            elems[t2f[e, j]].append(e)
    rowpts = [0]
    colind = []
    for i in range(nf):
        k = np.unique(t2f[elems[i]].ravel())
        colind.extend([i] + [int(v) for v in k if v != i])
        rowpts.append(len(colind))
    return np.array(rowpts), np.array(colind)


def global_assembly(AE, FE, t2f, elcon, rowpts, colind, ncu):
    """Hg (bsz, bsz, nblks), Rg (bsz, nrows) (globalassembly.m:11-39)."""
    ndf, ne = elcon.shape
    nfe = t2f.shape[1]
    npf = ndf // nfe
    bsz = ncu * npf
    He = AE.reshape((ncu, npf, nfe, ncu, npf, nfe, ne), order="F")
    Re = FE.reshape((ncu, npf, nfe, ne), order="F")
    el = elcon.reshape((npf, nfe, ne), order="F")
    nrows = len(rowpts) - 1
    Hg = np.zeros((bsz, bsz, len(colind)))
    Rg = np.zeros((bsz, nrows))
    for e in range(ne):
        for a in range(nfe):
            i = t2f[e, a]
            ind = el[:, a, e] - i * npf
            Rg[:, i] += Re[:, :, a, e][:, ind].reshape(bsz, order="F")
            for b in range(nfe):
                j = t2f[e, b]
                k = rowpts[i] + int(np.nonzero(colind[rowpts[i]:rowpts[i + 1]] == j)[0][0])
                jnd = el[:, b, e] - j * npf
                Hg[:, :, k] += He[:, :, a, :, :, b, e][:, ind][:, :, :, jnd].reshape((bsz, bsz), order="F")
    return Hg, Rg


def bilu0(Hg, rowpts, colind, order):
    """computeBILU0.m:3-45 with ordered2unordered = order."""
    nrows = len(rowpts) - 1
    unord2ord = np.empty(nrows, dtype=np.int64)
    unord2ord[np.asarray(order)] = np.arange(nrows)
    Mg = Hg.copy()
    for j in range(nrows):
        rj = order[j]
        nj = np.arange(rowpts[rj], rowpts[rj + 1])
        cj = colind[nj]
        jj = rowpts[rj]
        Mg[:, :, jj] = np.linalg.inv(Mg[:, :, jj])
        for ii_ in range(1, len(nj)):
            ri = cj[ii_]
            if unord2ord[ri] > j:
                ni = np.arange(rowpts[ri], rowpts[ri + 1])
                ci = colind[ni]
                aj = int(np.nonzero(ci == rj)[0][0])
                ij = rowpts[ri] + aj
                Mg[:, :, ij] = Mg[:, :, ij] @ Mg[:, :, jj]
                ii = rowpts[ri]
                ji = rowpts[rj] + ii_
                Mg[:, :, ii] = Mg[:, :, ii] - Mg[:, :, ij] @ Mg[:, :, ji]
                for l in range(1, len(ni)):
                    rij = ci[l]
                    if unord2ord[rij] > j:
                        for m in range(1, len(nj)):
                            if rij == cj[m]:
                                il = rowpts[ri] + l
                                jm = rowpts[rj] + m
                                Mg[:, :, il] = Mg[:, :, il] - Mg[:, :, ij] @ Mg[:, :, jm]
    return Mg


def bilu0_apply(Mg, rowpts, colind, order, x):
    """applyBILU0.m:10-52; x (bsz, nrows)."""
    nrows = len(rowpts) - 1
    unord2ord = np.empty(nrows, dtype=np.int64)
    unord2ord[np.asarray(order)] = np.arange(nrows)
    y = x.copy()
    for j in range(nrows):
        rj = order[j]
        y[:, rj] = x[:, rj]
        for i in range(1, rowpts[rj + 1] - rowpts[rj]):
            ri = colind[rowpts[rj] + i]
            if unord2ord[ri] < j:
                y[:, rj] = y[:, rj] - Mg[:, :, rowpts[rj] + i] @ y[:, ri]
    v = y.copy()
    for j in range(nrows):
        rj = order[nrows - j - 1]
        v[:, rj] = y[:, rj]
        for i in range(1, rowpts[rj + 1] - rowpts[rj]):
            ri = colind[rowpts[rj] + i]
            if unord2ord[ri] > unord2ord[rj]:
                v[:, rj] = v[:, rj] - Mg[:, :, rowpts[rj] + i] @ v[:, ri]
        v[:, rj] = Mg[:, :, rowpts[rj]] @ v[:, rj]
    return v


def cbsr_matvec(Hg, rowpts, colind, x):
    """matvec.m:3-12."""
    y = np.zeros_like(x)
    for i in range(len(rowpts) - 1):
        for k in range(rowpts[i], rowpts[i + 1]):
            y[:, i] += Hg[:, :, k] @ x[:, colind[k]]
    return y


def gmres_bilu0(matvec, precond, b, x=None, restart=20, tol=1e-6, maxit=None):
    """kernel/gmres_bilu0.m:1-103: left preconditioning, modified
    Gram-Schmidt.  b and x flat; returns (x, flag, iter)."""
    N = b.size
    x = np.zeros(N) if x is None else x.copy()
    maxit = N if maxit is None else maxit
    b = precond(b)
    nrmb = np.linalg.norm(b)
    flags, it = np.inf, 0
    while True:
        r = b - precond(matvec(x))
        beta = np.linalg.norm(r)
        V = np.zeros((N, restart + 1))
        hh = np.zeros((restart + 1, restart))
        V[:, 0] = r / beta
        y = np.zeros(0)
        for j in range(restart):
            w = precond(matvec(V[:, j]))
            for i in range(j + 1):
                hh[i, j] = V[:, i] @ w
                w = w - hh[i, j] * V[:, i]
            hh[j + 1, j] = np.linalg.norm(w)
            if hh[j + 1, j] != 0.0:
                V[:, j + 1] = w / hh[j + 1, j]
            else:
                break
            e1 = np.zeros(j + 2)
            e1[0] = beta
            y = np.linalg.lstsq(hh[:j + 2, :j + 1], e1, rcond=None)[0]
            res = np.linalg.norm(e1 - hh[:j + 2, :j + 1] @ y)
            it += 1
            if res / nrmb <= tol:
                flags = 0
                break
            if it >= maxit:
                flags = 1
                break
        x = x + V[:, :y.size] @ y
        if flags < np.inf:
            break
    return x, flags, it


def local_update(DUDG, DUDG_DUH, DUH, elcon):
    """DUDGT = DUDG + mapContractK(DUDG_DUH, DUH(:, elcon), ...) (hdg_solve.m:119-121).
    DUH (ncu, nsiz)."""
    npv, nco, ncu, ndf, ne = DUDG_DUH.shape
    DUHE = DUH[:, elcon]                                       # (ncu, ndf, ne)
    return DUDG + np.einsum("ackje,kje->ace", DUDG_DUH, DUHE)
